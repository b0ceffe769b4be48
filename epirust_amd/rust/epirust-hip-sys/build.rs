// Builds libepigpu.so (hipcc --offload-arch=gfx950, in-tree) and links it.
// EPIRUST_AMD_DIR points at the checkout that holds epirust_amd/; by default it is found
// three directories up from this crate (epirust_amd/rust/epirust-hip-sys).
use std::path::PathBuf;

fn main() {
    let root = match std::env::var("EPIRUST_AMD_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(std::env::var("CARGO_MANIFEST_DIR").unwrap()).join("../../.."),
    };
    let status = std::process::Command::new("python3")
        .args(["-m", "epirust_amd.build"])
        .current_dir(&root)
        .status()
        .expect("python3 with hipcc on PATH");
    assert!(status.success(), "hipcc build of libepigpu.so failed");
    let lib_dir = root.join("epirust_amd");
    println!("cargo:rustc-link-search=native={}", lib_dir.display());
    println!("cargo:rustc-link-lib=dylib=epigpu");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", lib_dir.display());
    println!("cargo:rerun-if-changed={}", lib_dir.join("csrc").display());
    println!("cargo:rerun-if-changed={}", root.join("include").join("epirust_gpu.h").display());
    println!("cargo:rerun-if-env-changed=EPIRUST_AMD_DIR");
}
