// rchem's build.rs with the MI355X library added (written for the rchem crate;
// not compiled in the image this repository was built in, which has no cargo).
//
// The reference build (build.rs:1-20) compiles libpyquante2 with cmake and
// binds cints.h with bindgen; that stays, so tho66.rs and its tests link as
// before.  With the `gpu` feature the library of this repository is built by
// its own Makefile (hipcc --offload-arch=gfx950) and rchem_gpu.h is bound the
// same way.  RCHEM_GPU_ROOT points at a checkout of this repository.
use bindgen;
use cmake;
use std::env;
use std::path::PathBuf;
use std::process::Command;

fn main() {
    let out_dir = PathBuf::from(env::var("OUT_DIR").unwrap());

    let dst = cmake::Config::new("libpyquante2").build();
    println!("cargo:rustc-link-search={}", dst.display());
    println!("cargo:rustc-link-lib=static=pyquante2");
    bindgen::Builder::default()
        .header("libpyquante2/cints.h")
        .parse_callbacks(Box::new(bindgen::CargoCallbacks::new()))
        .generate()
        .expect("Unable to generate bindings")
        .write_to_file(out_dir.join("bindings_libpyquante2.rs"))
        .expect("Couldn't write bindings!");

    if env::var_os("CARGO_FEATURE_GPU").is_some() {
        let root = PathBuf::from(env::var("RCHEM_GPU_ROOT").expect("RCHEM_GPU_ROOT"));
        let csrc = root.join("rchem_amd").join("csrc");
        let status = Command::new("make")
            .arg("-C")
            .arg(&csrc)
            .arg("-j8")
            .status()
            .expect("make (hipcc for gfx950)");
        assert!(status.success(), "building librchem_gpu.so failed");
        let lib_dir = root.join("rchem_amd");
        println!("cargo:rustc-link-search=native={}", lib_dir.display());
        println!("cargo:rustc-link-lib=dylib=rchem_gpu");
        println!("cargo:rustc-link-arg=-Wl,-rpath,{}", lib_dir.display());
        println!("cargo:rerun-if-changed={}", root.join("include").display());
        bindgen::Builder::default()
            .header(root.join("include").join("rchem_gpu.h").to_str().unwrap())
            .allowlist_function("rchem_.*")
            .allowlist_type("rchem_.*")
            .generate()
            .expect("bindgen rchem_gpu.h")
            .write_to_file(out_dir.join("bindings_rchem_gpu.rs"))
            .expect("Couldn't write bindings!");
    }
}
