//! Links libpar2hip.so.  PAR2HIP_LIB_DIR names the directory that holds it
//! (par2-rs_amd/ after `make -C par2-rs_amd/csrc`); the library itself
//! depends only on the ROCm runtime (libamdhip64.so).
fn main() {
    println!("cargo:rerun-if-env-changed=PAR2HIP_LIB_DIR");
    if let Ok(dir) = std::env::var("PAR2HIP_LIB_DIR") {
        println!("cargo:rustc-link-search=native={dir}");
    }
    println!("cargo:rustc-link-lib=dylib=par2hip");
}
