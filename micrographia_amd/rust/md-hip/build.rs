// Links librahman_hip.so (built by `python -m micrographia_amd.build`, hipcc for gfx950) from
// the directory named by RAHMAN_HIP_LIB, and records that directory as the run-time search path.
fn main() {
    println!("cargo:rerun-if-env-changed=RAHMAN_HIP_LIB");
    let dir = std::env::var("RAHMAN_HIP_LIB").expect("set RAHMAN_HIP_LIB to the directory holding librahman_hip.so");
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=rahman_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
}
