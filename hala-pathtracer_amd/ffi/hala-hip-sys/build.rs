// Links libhpt.so. HPT_LIB_DIR points at the directory holding it
// (hala-pathtracer_amd/lib after `make -C hala-pathtracer_amd/csrc`).
fn main() {
    let dir = std::env::var("HPT_LIB_DIR").unwrap_or_else(|_| "../../lib".to_string());
    println!("cargo:rustc-link-search=native={}", dir);
    println!("cargo:rustc-link-lib=dylib=hpt");
    println!("cargo:rerun-if-env-changed=HPT_LIB_DIR");
}
