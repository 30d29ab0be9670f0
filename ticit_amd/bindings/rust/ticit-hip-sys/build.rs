// Links libticit_hip.so. TICIT_HIP_LIB_DIR names the directory that holds it
// (the in-tree build writes ticit_amd/_lib/libticit_hip.so); without it the
// crate links against the library next to this crate's checkout.
use std::env;
use std::path::PathBuf;

fn main() {
    println!("cargo:rerun-if-env-changed=TICIT_HIP_LIB_DIR");
    let dir = match env::var("TICIT_HIP_LIB_DIR") {
        Ok(d) => PathBuf::from(d),
        Err(_) => PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap()).join("../../../_lib"),
    };
    println!("cargo:rustc-link-search=native={}", dir.display());
    println!("cargo:rustc-link-lib=dylib=ticit_hip");
    // the library resolves libamdhip64 from ROCm at run time
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", dir.display());
}
