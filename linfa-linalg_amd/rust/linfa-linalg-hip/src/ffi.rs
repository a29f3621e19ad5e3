//! Raw bindings of include/llh.h (the C ABI of libllh.so), f64 (`_d`) and f32
//! (`_s`).  Every function returns an `llh_status` (0 = OK) and fills `err`.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

pub type llh_status = i32;
pub const LLH_OK: llh_status = 0;
pub const LLH_ERR_NOT_SQUARE: llh_status = 1;
pub const LLH_ERR_NOT_POSITIVE_DEFINITE: llh_status = 2;
pub const LLH_ERR_WRONG_ROWS: llh_status = 3;
pub const LLH_ERR_INVALID_ARGUMENT: llh_status = 16;
pub const LLH_ERR_HIP: llh_status = 17;
pub const LLH_ERR_OUT_OF_MEMORY: llh_status = 18;
pub const LLH_ERR_NO_CONVERGENCE: llh_status = 19;
pub const LLH_ERR_REJECTED: llh_status = 20;

pub type llh_uplo = c_int;
pub const LLH_UPPER: llh_uplo = 0;
pub const LLH_LOWER: llh_uplo = 1;

pub type llh_precision = c_int;
pub const LLH_PREC_F64: llh_precision = 0;
pub const LLH_PREC_F32: llh_precision = 1;
pub const LLH_PREC_F32_REFINED: llh_precision = 2;
pub const LLH_PREC_F16: llh_precision = 3;
pub const LLH_FACTOR_FALLBACK_F64: i32 = 1;

pub const LLH_GEN_NORMAL: c_int = 0;
pub const LLH_GEN_UNIFORM: c_int = 1;

#[repr(C)]
#[derive(Clone, Copy)]
pub struct llh_error {
    pub code: i32,
    pub hip_error: i32,
    pub a: u64,
    pub b: u64,
    pub column: i64,
    pub message: [c_char; 192],
}

impl Default for llh_error {
    fn default() -> Self {
        llh_error { code: 0, hip_error: 0, a: 0, b: 0, column: -1, message: [0; 192] }
    }
}

#[repr(C)]
pub struct llh_context {
    _private: [u8; 0],
}
#[repr(C)]
pub struct llh_factor {
    _private: [u8; 0],
}
#[repr(C)]
pub struct llh_batch_plan {
    _private: [u8; 0],
}
#[repr(C)]
pub struct llh_server {
    _private: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct llh_kernel_stats {
    pub launches: u64,
    pub total_ms: f64,
    pub flops: f64,
    pub bytes: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct llh_refine_opts {
    pub max_iter: i32,
    pub flags: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct llh_refine_info {
    pub iterations: i32,
    pub converged: i32,
    pub fell_back: i32,
    pub precision_used: i32,
    pub residual: f64,
    pub anorm: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct llh_batch_config {
    pub small_limit: i32,
    pub small_waves: i32,
    pub ir_iterations: i32,
    pub medium_mode: i32,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct llh_server_config {
    pub hbm_budget_bytes: u64,
    pub streams: i32,
    pub max_inflight: i32,
    pub small_limit: i32,
    pub ir_iterations: i32,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct llh_request {
    pub matrix_id: u64,
    pub n: i32,
    pub m: i32,
    pub precision: i32,
    pub a_on_device: i32,
    pub a: *const f64,
    pub b: *const f64,
    pub x: *mut f64,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct llh_result {
    pub status: i32,
    pub cache_hit: i32,
    pub batched: i32,
    pub precision_used: i32,
    pub failed_row: i64,
    pub submit_s: f64,
    pub admit_s: f64,
    pub done_s: f64,
    pub residual: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct llh_server_stats {
    pub submitted: u64,
    pub completed: u64,
    pub batched: u64,
    pub cache_hits: u64,
    pub rejected: u64,
    pub steps: u64,
    pub bytes_in_use: u64,
    pub bytes_peak: u64,
    pub resident_factors: u64,
}

pub const LLH_SERVER_INJECT_STAGING: c_int = 1;
pub const LLH_SERVER_INJECT_EVENT: c_int = 2;
pub const LLH_SERVER_INJECT_SOLVE_SPIN: c_int = 3;

type Ctx = *mut llh_context;
type Err = *mut llh_error;

macro_rules! plain_and_dev {
    ($t:ty, $chol_dirty:ident, $chol:ident, $tri:ident, $is_tri:ident, $solve_tri:ident, $solve_sys:ident,
     $solvec:ident, $invc:ident, $dpotrf:ident, $dsolve_tri:ident, $dsolvec:ident, $dinvc:ident, $dortho:ident,
     $dapply:ident) => {
        extern "C" {
            pub fn $chol_dirty(ctx: Ctx, rows: usize, cols: usize, a: *mut $t, rs: isize, cs: isize,
                               err: Err) -> llh_status;
            pub fn $chol(ctx: Ctx, rows: usize, cols: usize, a: *mut $t, rs: isize, cs: isize, err: Err) -> llh_status;
            pub fn $tri(ctx: Ctx, rows: usize, cols: usize, a: *mut $t, rs: isize, cs: isize, uplo: llh_uplo,
                        err: Err) -> llh_status;
            pub fn $is_tri(ctx: Ctx, rows: usize, cols: usize, a: *const $t, rs: isize, cs: isize, uplo: llh_uplo,
                           out: *mut c_int, err: Err) -> llh_status;
            pub fn $solve_tri(ctx: Ctx, rows: usize, cols: usize, a: *const $t, rs: isize, cs: isize, b_rows: usize,
                              b_cols: usize, b: *mut $t, brs: isize, bcs: isize, uplo: llh_uplo, err: Err)
                              -> llh_status;
            pub fn $solve_sys(ctx: Ctx, rows: usize, cols: usize, a: *const $t, rs: isize, cs: isize,
                              diag: *const $t, dstride: isize, abs_diag: c_int, b_rows: usize, b_cols: usize,
                              b: *mut $t, brs: isize, bcs: isize, uplo: llh_uplo, err: Err) -> llh_status;
            pub fn $solvec(ctx: Ctx, rows: usize, cols: usize, a: *mut $t, rs: isize, cs: isize, b_rows: usize,
                           b_cols: usize, b: *mut $t, brs: isize, bcs: isize, err: Err) -> llh_status;
            pub fn $invc(ctx: Ctx, rows: usize, cols: usize, a: *mut $t, rs: isize, cs: isize, out: *mut $t,
                         ors: isize, ocs: isize, err: Err) -> llh_status;
            pub fn $dpotrf(ctx: Ctx, n: usize, a: *mut $t, lda: usize, zero_upper: c_int, err: Err) -> llh_status;
            pub fn $dsolve_tri(ctx: Ctx, uplo: llh_uplo, trans: c_int, n: usize, a: *const $t, lda: usize, m: usize,
                               b: *mut $t, ldb: usize, err: Err) -> llh_status;
            pub fn $dsolvec(ctx: Ctx, n: usize, a: *mut $t, lda: usize, m: usize, b: *mut $t, ldb: usize, err: Err)
                            -> llh_status;
            pub fn $dinvc(ctx: Ctx, n: usize, a: *mut $t, lda: usize, out: *mut $t, ldo: usize, err: Err)
                          -> llh_status;
            pub fn $dortho(ctx: Ctx, n: usize, k: usize, v: *mut $t, ldv: usize, r: *mut $t, ldr: usize, err: Err)
                           -> llh_status;
            pub fn $dapply(ctx: Ctx, n: usize, k: usize, v: *mut $t, ldv: usize, ky: usize, y: *const $t,
                           ldy: usize, l_yy: *const $t, ldl: usize, err: Err) -> llh_status;
        }
    };
}

plain_and_dev!(f64, llh_cholesky_inplace_dirty_d, llh_cholesky_inplace_d, llh_triangular_inplace_d,
               llh_is_triangular_d, llh_solve_triangular_inplace_d, llh_solve_triangular_system_d,
               llh_solvec_inplace_d, llh_invc_inplace_d, llh_dev_potrf_d, llh_dev_solve_triangular_d,
               llh_dev_solvec_d, llh_dev_invc_d, llh_dev_orthonormalize_d, llh_dev_apply_constraints_d);
plain_and_dev!(f32, llh_cholesky_inplace_dirty_s, llh_cholesky_inplace_s, llh_triangular_inplace_s,
               llh_is_triangular_s, llh_solve_triangular_inplace_s, llh_solve_triangular_system_s,
               llh_solvec_inplace_s, llh_invc_inplace_s, llh_dev_potrf_s, llh_dev_solve_triangular_s,
               llh_dev_solvec_s, llh_dev_invc_s, llh_dev_orthonormalize_s, llh_dev_apply_constraints_s);

extern "C" {
    pub fn llh_context_create(device: c_int, out: *mut Ctx, err: Err) -> llh_status;
    pub fn llh_context_destroy(ctx: Ctx);
    pub fn llh_context_set_stream(ctx: Ctx, hip_stream: *mut c_void, err: Err) -> llh_status;
    pub fn llh_context_stream(ctx: Ctx) -> *mut c_void;
    pub fn llh_status_string(code: i32) -> *const c_char;
    pub fn llh_context_set_profiling(ctx: Ctx, enable: c_int, err: Err) -> llh_status;
    pub fn llh_context_set_profiling_classes(ctx: Ctx, class_mask: u32, err: Err) -> llh_status;
    pub fn llh_context_kernel_stats(ctx: Ctx, kernel_class: c_int, out: *mut llh_kernel_stats, err: Err)
                                    -> llh_status;
    pub fn llh_context_set_solve_spin_limit(ctx: Ctx, limit: i64, err: Err) -> llh_status;

    pub fn llh_dev_solvec_noalloc_d(ctx: Ctx, n: usize, a: *mut f64, lda: usize, m: usize, b: *mut f64, ldb: usize,
                                    err: Err) -> llh_status;

    pub fn llh_dev_gen_matrix_d(ctx: Ctx, dist: c_int, seed: u64, stream: u64, rows: usize, cols: usize,
                                x: *mut f64, ldx: usize, err: Err) -> llh_status;
    pub fn llh_dev_gen_spd_d(ctx: Ctx, dist: c_int, seed: u64, n: usize, lam: f64, a: *mut f64, lda: usize,
                             b_work: *mut f64, err: Err) -> llh_status;

    pub fn llh_factor_create(ctx: Ctx, prec: llh_precision, n: usize, a: *const f64, lda: usize, flags: i32,
                             out: *mut *mut llh_factor, err: Err) -> llh_status;
    pub fn llh_factor_solve(ctx: Ctx, f: *mut llh_factor, m: usize, b: *const f64, ldb: usize, x: *mut f64,
                            ldx: usize, opts: *const llh_refine_opts, info: *mut llh_refine_info, err: Err)
                            -> llh_status;
    pub fn llh_factor_precision(f: *const llh_factor) -> llh_precision;
    pub fn llh_factor_device_bytes(f: *const llh_factor) -> usize;
    pub fn llh_factor_destroy(f: *mut llh_factor);
    pub fn llh_dev_solve_mixed(ctx: Ctx, prec: llh_precision, n: usize, a: *const f64, lda: usize, m: usize,
                               b: *const f64, ldb: usize, x: *mut f64, ldx: usize, opts: *const llh_refine_opts,
                               info: *mut llh_refine_info, err: Err) -> llh_status;

    pub fn llh_batch_plan_create(ctx: Ctx, prec: llh_precision, count: usize, n: *const i32, m: *const i32,
                                 config: *const llh_batch_config, out: *mut *mut llh_batch_plan, err: Err)
                                 -> llh_status;
    pub fn llh_batch_a_elems(plan: *const llh_batch_plan) -> usize;
    pub fn llh_batch_b_elems(plan: *const llh_batch_plan) -> usize;
    pub fn llh_batch_layout(plan: *const llh_batch_plan, a_off: *mut u64, lda: *mut i32, b_off: *mut u64,
                            err: Err) -> llh_status;
    pub fn llh_batch_get_config(plan: *const llh_batch_plan) -> llh_batch_config;
    pub fn llh_batch_solve(ctx: Ctx, plan: *mut llh_batch_plan, a: *mut c_void, b: *mut c_void, status: *mut i32,
                           err: Err) -> llh_status;
    pub fn llh_batch_autotune(ctx: Ctx, plan: *mut llh_batch_plan, a: *const c_void, b: *const c_void,
                              status: *mut i32, best_ms: *mut f64, err: Err) -> llh_status;
    pub fn llh_batch_plan_destroy(plan: *mut llh_batch_plan);

    pub fn llh_server_create(ctx: Ctx, config: *const llh_server_config, out: *mut *mut llh_server, err: Err)
                             -> llh_status;
    pub fn llh_server_submit(srv: *mut llh_server, req: *const llh_request, ticket: *mut u64, err: Err)
                             -> llh_status;
    pub fn llh_server_poll(srv: *mut llh_server, ticket: u64, out: *mut llh_result) -> c_int;
    pub fn llh_server_wait(srv: *mut llh_server, ticket: u64, out: *mut llh_result, err: Err) -> llh_status;
    pub fn llh_server_get_stats(srv: *mut llh_server, out: *mut llh_server_stats);
    pub fn llh_server_inject(srv: *mut llh_server, kind: c_int, value: i64, err: Err) -> llh_status;
    pub fn llh_server_evict_all(srv: *mut llh_server);
    pub fn llh_server_destroy(srv: *mut llh_server);
}

// The HIP runtime calls the device-resident buffer type needs (libamdhip64).
pub type hipError_t = c_int;
pub const HIP_MEMCPY_HOST_TO_DEVICE: c_int = 1;
pub const HIP_MEMCPY_DEVICE_TO_HOST: c_int = 2;
extern "C" {
    pub fn hipMalloc(ptr: *mut *mut c_void, size: usize) -> hipError_t;
    pub fn hipFree(ptr: *mut c_void) -> hipError_t;
    pub fn hipMemcpy2D(dst: *mut c_void, dpitch: usize, src: *const c_void, spitch: usize, width: usize,
                       height: usize, kind: c_int) -> hipError_t;
    pub fn hipDeviceSynchronize() -> hipError_t;
}
