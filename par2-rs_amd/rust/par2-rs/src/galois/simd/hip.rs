//! `GaloisSimdStrategy` on an MI355X (gfx950) through libpar2hip.
//!
//! File of the par2-rs crate: `src/galois/simd/hip.rs`, declared in
//! `src/galois/simd/mod.rs` as `#[cfg(feature = "hip")] pub mod hip;` and
//! registered first in `x86::register_strategies`
//! (`src/galois/simd/x86/mod.rs:13-26`), ahead of AVX2-Shuffle: both are
//! `Priority::Optimal` and `select_best` keeps the first registered of the
//! highest priority (`src/galois/simd/mod.rs:171-189`).
//!
//! The trait is infallible and the release profile aborts on panic
//! (`Cargo.toml:46`), so a non-zero status from the library falls back to the
//! scalar strategy for that call.  The trait is called per slice or per
//! 128 KiB region, where the PCIe round trip would dominate, so calls below
//! `GPU_MIN_WORDS` stay on the CPU; the coarse hooks in the encoder and the
//! repairer (whole sets through `par2hip_encode` / `par2hip_reconstruct` /
//! `par2hip_repair_files`) are where the GPU pays.

use super::{GaloisSimdStrategy, Priority};
use crate::galois::scalar;
use par2hip_sys as ffi;

/// One context on device 0.  libpar2hip serialises the calls made on one
/// context, so the strategy can be shared by the Rayon workers.
pub struct HipStrategy {
    ctx: *mut ffi::par2hip_ctx,
}

// SAFETY: the context is only handed to libpar2hip, which locks it per call.
unsafe impl Send for HipStrategy {}
unsafe impl Sync for HipStrategy {}

/// Below this many u16 words per call the CPU strategy is faster than the
/// host-device round trip.
const GPU_MIN_WORDS: usize = 1 << 20;

impl HipStrategy {
    /// A strategy on device 0, or None when no gfx950 device is usable.
    pub fn new() -> Option<Self> {
        let mut ctx = core::ptr::null_mut();
        // SAFETY: plain out-pointer; the library reports failure as a status.
        let rc = unsafe { ffi::par2hip_ctx_create(0, &mut ctx) };
        (rc == ffi::PAR2HIP_OK).then_some(HipStrategy { ctx })
    }
}

impl Drop for HipStrategy {
    fn drop(&mut self) {
        // SAFETY: created by par2hip_ctx_create and not used after this.
        unsafe {
            ffi::par2hip_ctx_destroy(self.ctx);
        }
    }
}

impl GaloisSimdStrategy for HipStrategy {
    fn name(&self) -> &'static str {
        "HIP-gfx950"
    }

    fn is_available(&self) -> bool {
        // SAFETY: no arguments; the device count is probed once and cached.
        unsafe { ffi::par2hip_available() > 0 }
    }

    fn priority(&self) -> Priority {
        Priority::Optimal
    }

    unsafe fn mul_slice(&self, scalar: u16, data: &mut [u16]) {
        if data.len() < GPU_MIN_WORDS
            || ffi::par2hip_gf_mul_slice(self.ctx, scalar, data.as_mut_ptr(), data.len()) != ffi::PAR2HIP_OK
        {
            scalar::gf_mul_slice_scalar(scalar, data);
        }
    }

    unsafe fn muladd(&self, dst: &mut [u16], src: &[u16], scalar: u16) {
        let len = dst.len().min(src.len());
        if len < GPU_MIN_WORDS
            || ffi::par2hip_gf_muladd(self.ctx, dst.as_mut_ptr(), src.as_ptr(), len, scalar) != ffi::PAR2HIP_OK
        {
            scalar::gf_muladd_scalar(dst, src, scalar);
        }
    }

    unsafe fn muladd_region(
        &self,
        destinations: &mut [&mut [u16]],
        sources: &[&[u16]],
        coefficients: &[&[u16]],
        region_offset: usize,
        region_size: usize,
    ) {
        // every row and source must cover the region, and there must be a
        // coefficient row per destination; otherwise the scalar strategy
        // gives the reference's behaviour (including its panics)
        let end = region_offset.checked_add(region_size);
        let covered = |len: usize| end.map_or(false, |e| e <= len);
        let big = region_size.saturating_mul(sources.len()).saturating_mul(destinations.len()) >= GPU_MIN_WORDS;
        let fits = coefficients.len() >= destinations.len()
            && destinations.iter().all(|d| covered(d.len()))
            && sources.iter().all(|s| covered(s.len()));
        if big && fits {
            let dsts: Vec<*mut u16> = destinations.iter_mut().map(|d| d.as_mut_ptr()).collect();
            let srcs: Vec<*const u16> = sources.iter().map(|s| s.as_ptr()).collect();
            let rows: Vec<*const u16> = coefficients[..dsts.len()].iter().map(|c| c.as_ptr()).collect();
            // a row shorter than the source list applies only its own
            // entries (`.enumerate().take(num_srcs)`, src/galois/scalar.rs:100)
            let lens: Vec<usize> = coefficients[..dsts.len()].iter().map(|c| c.len()).collect();
            let rc = ffi::par2hip_gf_muladd_region(
                self.ctx,
                dsts.as_ptr(),
                dsts.len(),
                srcs.as_ptr(),
                srcs.len(),
                rows.as_ptr(),
                lens.as_ptr(),
                region_offset,
                region_size,
            );
            if rc == ffi::PAR2HIP_OK {
                return;
            }
        }
        scalar::gf_muladd_region_scalar(destinations, sources, coefficients, region_offset, region_size);
    }

    // supports_shuffle2x() keeps its default `false`: the GPU consumes
    // interleaved little-endian words, so the shuffle2x hooks are the trait's
    // default no-op conversions and plain muladd calls.
}
