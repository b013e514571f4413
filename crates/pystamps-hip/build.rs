// Links libpystamps_hip.so from the directory named by PYSTAMPS_HIP_LIB_DIR
// (the library is built by `python -m pystamps_amd.build`).
fn main() {
    let dir = std::env::var("PYSTAMPS_HIP_LIB_DIR")
        .expect("PYSTAMPS_HIP_LIB_DIR: directory of libpystamps_hip.so");
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=pystamps_hip");
    println!("cargo:rerun-if-env-changed=PYSTAMPS_HIP_LIB_DIR");
}
