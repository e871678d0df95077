// gf_mfma.hip -- placeholder translation unit for the bit-sliced MFMA
// contraction (filled in once the LDS-table variant is parity-green).
#include "common.hpp"
