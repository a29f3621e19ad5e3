//! Builds libllh.so (hand-written HIP kernels for gfx950 and the C ABI of
//! include/llh.h) with the library's own Makefile, and links it.
use std::env;
use std::path::PathBuf;
use std::process::Command;

fn main() {
    let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
    // linfa-linalg_amd/rust/linfa-linalg-hip -> linfa-linalg_amd
    let pkg = manifest.parent().and_then(|p| p.parent()).expect("crate lives under linfa-linalg_amd/rust/");
    let hipcc = env::var("HIPCC").unwrap_or_else(|_| "/opt/rocm/bin/hipcc".to_string());
    let arch = env::var("LLH_OFFLOAD_ARCH").unwrap_or_else(|_| "gfx950".to_string());
    let status = Command::new("make")
        .arg("-C")
        .arg(pkg)
        .arg(format!("HIPCC={}", hipcc))
        .arg(format!("ARCH={}", arch))
        .status()
        .expect("running make for libllh.so");
    assert!(status.success(), "building libllh.so failed");
    let libdir = pkg.join("linfa_linalg_hip");
    println!("cargo:rustc-link-search=native={}", libdir.display());
    println!("cargo:rustc-link-lib=dylib=llh");
    // the HIP runtime, for the device-resident buffer type (hipMalloc / hipMemcpy2D)
    let rocm = env::var("ROCM_PATH").unwrap_or_else(|_| "/opt/rocm".to_string());
    println!("cargo:rustc-link-search=native={}/lib", rocm);
    println!("cargo:rustc-link-lib=dylib=amdhip64");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", libdir.display());
    println!("cargo:rerun-if-changed={}", pkg.join("csrc").display());
    println!("cargo:rerun-if-changed={}", pkg.join("../include/llh.h").display());
    println!("cargo:rerun-if-env-changed=HIPCC");
    println!("cargo:rerun-if-env-changed=LLH_OFFLOAD_ARCH");
}
