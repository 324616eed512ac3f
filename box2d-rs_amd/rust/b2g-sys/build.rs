// libb2g.so is built by `make -C box2d-rs_amd` (hipcc --offload-arch=gfx950) into box2d-rs_amd/b2g/.
fn main() {
    let dir = std::env::var("B2G_LIB_DIR").unwrap_or_else(|_| "../../b2g".to_string());
    println!("cargo:rustc-link-search=native={}", dir);
    println!("cargo:rustc-link-lib=dylib=b2g");
    println!("cargo:rerun-if-env-changed=B2G_LIB_DIR");
}
