//! Links libvoronoid_hip.so. VORONOID_HIP_LIB_DIR names the directory that holds it (the
//! in-tree build puts it next to vorothree_amd/build.py); the HIP runtime it depends on is found
//! through the loader's usual paths (ROCm's lib directory).
fn main() {
    let dir = std::env::var("VORONOID_HIP_LIB_DIR").unwrap_or_else(|_| {
        let here = std::env::var("CARGO_MANIFEST_DIR").expect("cargo sets CARGO_MANIFEST_DIR");
        format!("{here}/..")
    });
    println!("cargo:rustc-link-search=native={dir}");
    println!("cargo:rustc-link-lib=dylib=voronoid_hip");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{dir}");
    println!("cargo:rerun-if-env-changed=VORONOID_HIP_LIB_DIR");
}
