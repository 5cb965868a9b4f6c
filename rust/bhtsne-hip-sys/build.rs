//! Links libbhtsne_hip.so.  BHTSNE_HIP_LIB_DIR names the directory that holds it (the library is
//! built by `python -m bhtsne_amd.build`, which writes bhtsne_amd/libbhtsne_hip.so).
fn main() {
    println!("cargo:rerun-if-env-changed=BHTSNE_HIP_LIB_DIR");
    let dir = std::env::var("BHTSNE_HIP_LIB_DIR")
        .expect("set BHTSNE_HIP_LIB_DIR to the directory of libbhtsne_hip.so");
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=bhtsne_hip");
    // the library finds its HIP runtime through its own RUNPATH; the binary needs the library's
    println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
}
