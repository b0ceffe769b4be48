// Builds libepigpu.so (hipcc --offload-arch=gfx950, in-tree) and links it.
// EPIRUST_AMD_DIR points at the checkout that holds epirust_amd/; by default it is found
// three directories up from this crate (epirust_amd/rust/epirust-hip-sys).
// EPIRUST_DISEASE_HANDLER=HEADER:TYPE builds and links a library with that disease handler
// compiled in instead of the reference's Disease (the engine's DiseaseHandler type parameter).
use std::path::PathBuf;

fn main() {
    let root = match std::env::var("EPIRUST_AMD_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(std::env::var("CARGO_MANIFEST_DIR").unwrap()).join("../../.."),
    };
    let mut args = vec!["-m".to_string(), "epirust_amd.build".to_string()];
    let mut lib_dir = root.join("epirust_amd");
    if let Ok(handler) = std::env::var("EPIRUST_DISEASE_HANDLER") {
        let out_dir = PathBuf::from(std::env::var("OUT_DIR").unwrap());
        args.extend(["--disease-handler".to_string(), handler,
                     "--out".to_string(), out_dir.join("libepigpu.so").display().to_string()]);
        lib_dir = out_dir;
    }
    let status = std::process::Command::new("python3")
        .args(&args)
        .current_dir(&root)
        .status()
        .expect("python3 with hipcc on PATH");
    assert!(status.success(), "hipcc build of libepigpu.so failed");
    println!("cargo:rustc-link-search=native={}", lib_dir.display());
    println!("cargo:rustc-link-lib=dylib=epigpu");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", lib_dir.display());
    println!("cargo:rerun-if-changed={}", root.join("epirust_amd").join("csrc").display());
    println!("cargo:rerun-if-changed={}", root.join("include").join("epirust_gpu.h").display());
    println!("cargo:rerun-if-env-changed=EPIRUST_AMD_DIR");
    println!("cargo:rerun-if-env-changed=EPIRUST_DISEASE_HANDLER");
}
