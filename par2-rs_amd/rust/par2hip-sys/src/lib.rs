//! Raw FFI declarations of `include/par2hip.h`, one for one: every constant,
//! every `#[repr(C)]` struct field for field, every entry point with the
//! header's argument types.  The safe wrapper par2-rs uses is
//! `src/galois/simd/hip.rs` (the `GaloisSimdStrategy` implementation) plus
//! the coarse hooks in the encoder and repairer.
#![allow(non_camel_case_types)]
#![no_std]

use core::ffi::{c_char, c_void};

// ---- status codes (int32_t; 0 = success)
pub const PAR2HIP_OK: i32 = 0;
pub const PAR2HIP_EINVAL: i32 = -1;
pub const PAR2HIP_ENODEV: i32 = -2;
pub const PAR2HIP_ENOMEM: i32 = -3;
pub const PAR2HIP_ESINGULAR: i32 = -4;
pub const PAR2HIP_EHIP: i32 = -5;
pub const PAR2HIP_ERECOVERY: i32 = -6;
pub const PAR2HIP_EREPAIR: i32 = -7;
pub const PAR2HIP_ECANCELLED: i32 = -8;
pub const PAR2HIP_ETIMEDOUT: i32 = -9;
pub const PAR2HIP_PRIORITY_OPTIMAL: i32 = 40;
pub const PAR2HIP_JOB_ENCODE: i32 = 0;
pub const PAR2HIP_JOB_REPAIR: i32 = 1;
pub const PAR2HIP_VOLUMES_SINGLE: i32 = 0;
pub const PAR2HIP_VOLUMES_EXPONENTIAL: i32 = 1;

// ---- opaque handles
#[repr(C)]
pub struct par2hip_ctx {
    _p: [u8; 0],
}
#[repr(C)]
pub struct par2hip_batch_plan {
    _p: [u8; 0],
}
#[repr(C)]
pub struct par2hip_server {
    _p: [u8; 0],
}

// ---- structs
#[repr(C)]
#[derive(Clone, Copy)]
pub struct par2hip_device_info {
    pub compute_units: i32,
    pub hash_cus: i32,
    pub arch: [c_char; 32],
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct par2hip_job {
    pub kind: i32,
    pub reserved: i32,
    pub d_slices: *mut c_void,
    pub n: u64,
    pub slice_size: u64,
    pub exps: *const u32,
    pub count: u64,
    pub d_recovery: *mut c_void,
    pub missing: *const u32,
    pub d_out: *mut c_void,
    pub d_md5: *mut c_void,
    pub d_crc: *mut c_void,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct par2hip_batch_plan_tuning {
    pub policy: i32,
    pub policies: i32,
    pub ms: [f64; 4],
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct par2hip_request {
    pub job: par2hip_job,
    pub hashes: i32,
    pub reserved: i32,
    pub arrival_ms: f64,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct par2hip_result {
    pub status: i32,
    pub batch_size: i32,
    pub d_out: *mut c_void,
    pub d_md5: *mut c_void,
    pub d_crc: *mut c_void,
    pub t_arrival_ms: f64,
    pub t_submit_ms: f64,
    pub t_admit_ms: f64,
    pub t_start_ms: f64,
    pub t_done_ms: f64,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct par2hip_server_stats {
    pub submitted: u64,
    pub completed: u64,
    pub failed: u64,
    pub batches: u64,
    pub max_batch_seen: u64,
    pub charged: u64,
    pub peak_charged: u64,
    pub pending: u64,
    pub admitted: u64,
    pub arena_fallbacks: u64,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct par2hip_input_file {
    pub name: *const c_char,
    pub data: *const c_void,
    pub len: u64,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct par2hip_output_file {
    pub path: [c_char; 1024],
    pub bytes: u64,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct par2hip_set_info {
    pub block_size: u64,
    pub blocks: u64,
    pub files: u32,
    pub recovery_slices: u32,
    pub min_exponent: u32,
    pub max_exponent: u32,
    pub set_id: [u8; 16],
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct par2hip_repair_report {
    pub block_size: u64,
    pub blocks: u64,
    pub blocks_damaged: u64,
    pub files: u32,
    pub files_damaged: u32,
    pub files_missing: u32,
    pub recovery_slices: u32,
    pub repaired: i32,
    pub reserved: i32,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct par2hip_kernel_stat {
    pub name: [c_char; 48],
    pub launches: u64,
    pub total_ms: f64,
}

extern "C" {
    // ---- device
    pub fn par2hip_available() -> i32;
    pub fn par2hip_strategy_name() -> *const c_char;
    pub fn par2hip_strategy_priority() -> i32;
    pub fn par2hip_ctx_create(device: i32, out: *mut *mut par2hip_ctx) -> i32;
    pub fn par2hip_ctx_destroy(ctx: *mut par2hip_ctx) -> i32;
    pub fn par2hip_last_error(ctx: *mut par2hip_ctx) -> *const c_char;
    pub fn par2hip_get_device_info(ctx: *mut par2hip_ctx, out: *mut par2hip_device_info) -> i32;
    pub fn par2hip_dev_alloc(ctx: *mut par2hip_ctx, bytes: usize, dptr: *mut *mut c_void) -> i32;
    pub fn par2hip_dev_free(ctx: *mut par2hip_ctx, dptr: *mut c_void) -> i32;
    pub fn par2hip_h2d(ctx: *mut par2hip_ctx, dst: *mut c_void, src: *const c_void, bytes: usize) -> i32;
    pub fn par2hip_d2h(ctx: *mut par2hip_ctx, dst: *mut c_void, src: *const c_void, bytes: usize) -> i32;
    pub fn par2hip_d2d(ctx: *mut par2hip_ctx, dst: *mut c_void, src: *const c_void, bytes: usize) -> i32;
    pub fn par2hip_sync(ctx: *mut par2hip_ctx) -> i32;

    // ---- strategy-shaped (GaloisSimdStrategy, src/galois/simd/mod.rs:24-128)
    pub fn par2hip_gf_mul_slice(ctx: *mut par2hip_ctx, scalar: u16, data: *mut u16, len: usize) -> i32;
    pub fn par2hip_gf_muladd(ctx: *mut par2hip_ctx, dst: *mut u16, src: *const u16, len: usize, scalar: u16) -> i32;
    pub fn par2hip_gf_muladd_region(
        ctx: *mut par2hip_ctx,
        dsts: *const *mut u16,
        ndst: usize,
        srcs: *const *const u16,
        nsrc: usize,
        coeff_rows: *const *const u16,
        coeff_lens: *const usize,
        region_offset: usize,
        region_size: usize,
    ) -> i32;

    // ---- encode (StreamingEncoder, src/par2_rs.rs:687-902)
    pub fn par2hip_encode(
        ctx: *mut par2hip_ctx,
        slices: *const u8,
        n: usize,
        slice_size: usize,
        exps: *const u32,
        r: usize,
        recovery: *mut u8,
        md5_out: *mut u8,
        crc_out: *mut u32,
    ) -> i32;
    pub fn par2hip_encode_dev(
        ctx: *mut par2hip_ctx,
        d_slices: *const c_void,
        n: usize,
        slice_size: usize,
        exps: *const u32,
        r: usize,
        d_recovery: *mut c_void,
        d_md5: *mut c_void,
        d_crc: *mut c_void,
    ) -> i32;
    pub fn par2hip_encode_dev_noalloc(
        ctx: *mut par2hip_ctx,
        d_slices: *const c_void,
        n: usize,
        slice_size: usize,
        exps: *const u32,
        r: usize,
        d_recovery: *mut c_void,
        d_md5: *mut c_void,
        d_crc: *mut c_void,
    ) -> i32;

    // ---- repair (src/par2_rs.rs:319-675, src/repair.rs:536-970)
    pub fn par2hip_gf_invert(ctx: *mut par2hip_ctx, a: *const u16, m: usize, inv: *mut u16) -> i32;
    pub fn par2hip_reconstruct(
        ctx: *mut par2hip_ctx,
        slices: *const u8,
        n: usize,
        slice_size: usize,
        missing: *const u32,
        m: usize,
        recovery: *const u8,
        rec_exps: *const u32,
        out: *mut u8,
    ) -> i32;
    pub fn par2hip_reconstruct_dev(
        ctx: *mut par2hip_ctx,
        d_slices: *mut c_void,
        n: usize,
        slice_size: usize,
        missing: *const u32,
        m: usize,
        d_recovery: *const c_void,
        rec_exps: *const u32,
        d_out: *mut c_void,
        out_in_place: i32,
    ) -> i32;
    pub fn par2hip_repair_set_dev(
        ctx: *mut par2hip_ctx,
        d_slices: *mut c_void,
        n: usize,
        slice_size: usize,
        expected_md5: *const u8,
        expected_crc: *const u32,
        d_recovery: *const c_void,
        rec_exps: *const u32,
        nrec: usize,
        d_out: *mut c_void,
        damaged_out: *mut u32,
        ndamaged: *mut usize,
    ) -> i32;

    // ---- batched
    pub fn par2hip_batch_run_dev(ctx: *mut par2hip_ctx, jobs: *const par2hip_job, njobs: usize) -> i32;
    pub fn par2hip_batch_plan_create(
        ctx: *mut par2hip_ctx,
        jobs: *const par2hip_job,
        njobs: usize,
        use_graph: i32,
        out: *mut *mut par2hip_batch_plan,
    ) -> i32;
    pub fn par2hip_batch_plan_launch(plan: *mut par2hip_batch_plan) -> i32;
    pub fn par2hip_batch_plan_destroy(plan: *mut par2hip_batch_plan) -> i32;
    pub fn par2hip_batch_plan_tuning_info(plan: *mut par2hip_batch_plan, out: *mut par2hip_batch_plan_tuning) -> i32;

    // ---- serving
    pub fn par2hip_server_create(
        device: i32,
        hbm_budget: u64,
        lanes: i32,
        max_batch: i32,
        out: *mut *mut par2hip_server,
    ) -> i32;
    pub fn par2hip_server_destroy(srv: *mut par2hip_server) -> i32;
    pub fn par2hip_server_now_ms(srv: *mut par2hip_server) -> f64;
    pub fn par2hip_server_submit(srv: *mut par2hip_server, req: *const par2hip_request, ticket: *mut u64) -> i32;
    pub fn par2hip_server_wait(srv: *mut par2hip_server, ticket: u64, res: *mut par2hip_result) -> i32;
    pub fn par2hip_server_wait_timeout(
        srv: *mut par2hip_server,
        ticket: u64,
        timeout_ms: f64,
        res: *mut par2hip_result,
    ) -> i32;
    pub fn par2hip_server_release(srv: *mut par2hip_server, ticket: u64) -> i32;
    pub fn par2hip_server_get_stats(srv: *mut par2hip_server, st: *mut par2hip_server_stats) -> i32;

    // ---- .par2 files (Par2Creator, Par2Info, Par2Repairer)
    pub fn par2hip_create_par2(
        ctx: *mut par2hip_ctx,
        files: *const par2hip_input_file,
        nfiles: usize,
        block_size: u64,
        redundancy_percent: f32,
        scheme: i32,
        out_stem: *const c_char,
        outs: *mut par2hip_output_file,
        max_outs: usize,
        nouts: *mut usize,
    ) -> i32;
    pub fn par2hip_par2_info(par2_path: *const c_char, info: *mut par2hip_set_info) -> i32;
    pub fn par2hip_repair_files(
        ctx: *mut par2hip_ctx,
        par2_path: *const c_char,
        base_dir: *const c_char,
        verify_only: i32,
        report: *mut par2hip_repair_report,
    ) -> i32;
    pub fn par2hip_repair_files_cancellable(
        ctx: *mut par2hip_ctx,
        par2_path: *const c_char,
        base_dir: *const c_char,
        verify_only: i32,
        cancel: *const u8,
        report: *mut par2hip_repair_report,
    ) -> i32;

    // ---- slice hashes (IFSC)
    pub fn par2hip_hash_slices(
        ctx: *mut par2hip_ctx,
        slices: *const u8,
        n: usize,
        slice_size: usize,
        slice_stride: usize,
        lens: *const u64,
        md5_out: *mut u8,
        crc_out: *mut u32,
    ) -> i32;
    pub fn par2hip_hash_slices_dev(
        ctx: *mut par2hip_ctx,
        d_slices: *const c_void,
        n: usize,
        slice_size: usize,
        slice_stride: usize,
        lens: *const u64,
        d_md5: *mut c_void,
        d_crc: *mut c_void,
    ) -> i32;

    // ---- tuning and profiling
    pub fn par2hip_set_gf_kernel(ctx: *mut par2hip_ctx, which: i32) -> i32;
    pub fn par2hip_set_autotune(ctx: *mut par2hip_ctx, mode: i32) -> i32;
    pub fn par2hip_force_tile_shape(ctx: *mut par2hip_ctx, v: i32, g: i32, d: i32) -> i32;
    pub fn par2hip_autotune_save(path: *const c_char) -> i32;
    pub fn par2hip_autotune_load(path: *const c_char, loaded: *mut usize) -> i32;
    pub fn par2hip_profile_enable(ctx: *mut par2hip_ctx, on: i32) -> i32;
    pub fn par2hip_profile_read(
        ctx: *mut par2hip_ctx,
        out: *mut par2hip_kernel_stat,
        max: usize,
        count: *mut usize,
    ) -> i32;
}
