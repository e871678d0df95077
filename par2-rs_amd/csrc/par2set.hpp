// par2set.hpp -- loading a PAR2 set on the host: the packet parser, volume
// discovery, canonical file order and the block layout of the protected
// files.  Host-only C++ (no HIP): it reads untrusted files, so it is also
// built on its own under AddressSanitizer/UBSan by the CPU tests.
//
//   parse   src/parser.rs:181-547 (packets, resync, MD5 check of every packet
//           but the recovery slices, which are located and read at repair
//           time; volumes = every .par2 beside the index with the same
//           recovery set ID), canonical order src/parser.rs:556-596
//   layout  src/repair.rs:499-545 (div_ceil / checked_add / try_from there;
//           checked arithmetic and explicit limits here)
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/par2hip.h"

namespace par2hip {

using Id = std::array<uint8_t, 16>;

struct FileRec {
  Id id{}, md5{}, md5k{};
  uint64_t length = 0;
  std::string name;
};
struct RecSlice {
  uint32_t exponent;
  std::string path;
  uint64_t data_offset;  // of the recovery bytes (after the exponent)
};
struct Par2Set {
  bool have_set_id = false;
  Id set_id{};
  uint64_t block_size = 0;
  std::vector<Id> main_order;
  std::vector<FileRec> files_in_order;  // FileDesc order of first appearance
  std::map<Id, FileRec> files;
  std::map<Id, std::vector<std::pair<Id, uint32_t>>> ifsc;  // md5, crc32 per slice
  std::vector<RecSlice> recs;
};

// One .par2 file (parser.rs:181-400).  recovery_only: a volume beside the
// index (only its recovery slices are taken, parser.rs:443-520).
int32_t parse_par2_file(const std::string &path, Par2Set &s, bool recovery_only, std::string *err);
// The index and every .par2 beside it with the same recovery set ID.
// PAR2HIP_EINVAL when the index holds no packet or no usable block size.
int32_t load_par2_set(const std::string &index, Par2Set &s, std::string *err);
// Files in the canonical block order (parser.rs:566-596): the Main packet's
// list first, then any other described file by File ID (byte 15 first).
std::vector<FileRec> canonical_files(const Par2Set &s);
// A stored file name the repairer may open, write and rename under the base
// directory: relative, non-empty, no ".." component.
bool safe_name(const std::string &name);
// Sorted by exponent, one slice per exponent (the first found).
std::vector<RecSlice> recovery_by_exponent(const Par2Set &s);
// Directory part of a path ("." when there is none).
std::string dir_of(const std::string &p);

// Blocks of the files in canonical order: file i covers global blocks
// [first[i], first[i] + nblk[i]); n blocks of block_size bytes in all.
struct BlockLayout {
  std::vector<uint64_t> first, nblk;
  uint64_t n = 0;
  uint64_t bytes = 0;  // n * block_size
};
// Layout with every step checked: PAR2HIP_EINVAL (and nothing else done)
// when the block count passes max_blocks, or n * block_size passes
// max_bytes (the caller's memory limit); block_size must be non-zero.
int32_t block_layout(const std::vector<FileRec> &files, uint64_t block_size, uint64_t max_blocks,
                     uint64_t max_bytes, BlockLayout *out, std::string *err);
// ceil(len / bs) without overflow (bs > 0).
inline uint64_t blocks_of(uint64_t len, uint64_t bs) { return len / bs + (len % bs != 0); }

// Block size of a new set (Par2Creator's automatic size): the reference's
// calculate_block_size (src/creator.rs:29-67: ~2000 blocks, 2 KiB..4 MiB,
// a multiple of 4, at most 65535 blocks).  Where that size would give more
// than max_blocks input blocks over the files (32768 distinct PAR2 constants;
// sets over ~128 GiB), the smallest multiple of 4 that keeps the count within
// it is taken instead, so large sets are created rather than refused.
// Returns 0 when no size can (max_blocks or more non-empty files).
uint64_t create_block_size(const uint64_t *file_lens, size_t nfiles, uint64_t max_blocks);
// The reference's rule alone (creator.rs:29-67) for a total size.
uint64_t reference_block_size(uint64_t total);

}  // namespace par2hip
