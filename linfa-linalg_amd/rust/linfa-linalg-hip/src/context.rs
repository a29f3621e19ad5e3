//! One library context (one HIP stream) per thread, created on first use on
//! the current device; calls on a context never overlap, as the C ABI requires.
use std::cell::RefCell;
use std::ffi::CStr;

use crate::ffi;
use linfa_linalg::{LinalgError, Result};

struct Context(*mut ffi::llh_context);

impl Drop for Context {
    fn drop(&mut self) {
        unsafe { ffi::llh_context_destroy(self.0) }
    }
}

thread_local! {
    static CTX: RefCell<Option<Context>> = RefCell::new(None);
}

fn message(e: &ffi::llh_error) -> String {
    unsafe { CStr::from_ptr(e.message.as_ptr()) }.to_string_lossy().into_owned()
}

/// Runs one C-ABI call with this thread's context and maps its status to the
/// reference's errors (lib.rs:33-60); device failures panic.
pub(crate) fn call(f: impl FnOnce(*mut ffi::llh_context, *mut ffi::llh_error) -> ffi::llh_status) -> Result<()> {
    CTX.with(|cell| {
        let mut slot = cell.borrow_mut();
        if slot.is_none() {
            let mut ctx: *mut ffi::llh_context = std::ptr::null_mut();
            let mut err = ffi::llh_error::default();
            let st = unsafe { ffi::llh_context_create(-1, &mut ctx, &mut err) };
            if st != ffi::LLH_OK {
                panic!("linfa-linalg-hip: context creation failed: {}", message(&err));
            }
            *slot = Some(Context(ctx));
        }
        let ctx = slot.as_ref().map(|c| c.0).unwrap();
        let mut err = ffi::llh_error::default();
        let st = f(ctx, &mut err);
        to_result(st, &err)
    })
}

pub(crate) fn to_result(st: ffi::llh_status, err: &ffi::llh_error) -> Result<()> {
    match st {
        ffi::LLH_OK => Ok(()),
        ffi::LLH_ERR_NOT_SQUARE => Err(LinalgError::NotSquare { rows: err.a as usize, cols: err.b as usize }),
        ffi::LLH_ERR_NOT_POSITIVE_DEFINITE => Err(LinalgError::NotPositiveDefinite),
        ffi::LLH_ERR_WRONG_ROWS => Err(LinalgError::WrongRows { expected: err.a as usize, actual: err.b as usize }),
        other => {
            let name = unsafe { CStr::from_ptr(ffi::llh_status_string(other)) }.to_string_lossy().into_owned();
            panic!("linfa-linalg-hip: {} ({})", name, message(err))
        }
    }
}
