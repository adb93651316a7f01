// build.rs of the reference crate's `hip` feature: compiles the HIP sources of
// gstools-core_amd/csrc for gfx950 with hipcc (the flags of
// gstools-core_amd/build_native.py) into libgstools_core_amd.so in OUT_DIR and
// links the crate against it.  Copy next to the crate's Cargo.toml; set
// GSC_AMD_SRC to the gstools-core_amd directory (default: ../gstools-core_amd)
// and HIPCC if hipcc is not /opt/rocm/bin/hipcc.  Without the `hip` feature
// it does nothing, so the CPU build of the crate is unchanged.
use std::env;
use std::path::{Path, PathBuf};
use std::process::Command;

fn hipcc() -> String {
    env::var("HIPCC").unwrap_or_else(|_| "/opt/rocm/bin/hipcc".to_string())
}

fn compile(src: &Path, include: &Path, csrc: &Path, obj: &Path) {
    let name = src.file_name().unwrap().to_string_lossy().to_string();
    let mut cmd = Command::new(hipcc());
    cmd.args(["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950"])
        // no packed FP32 instructions anywhere in the device code
        .args(["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"])
        .arg("-munsafe-fp-atomics")
        .arg("-I").arg(include)
        .arg("-I").arg(csrc);
    if name == "direct.hip" || name == "direct_local.hip" {
        // MFMA results in arch VGPRs for the direct kernels
        cmd.args(["-mllvm", "-amdgpu-mfma-vgpr-form=1"]);
    }
    cmd.arg("-c").arg(src).arg("-o").arg(obj);
    let st = cmd.status().expect("running hipcc");
    assert!(st.success(), "hipcc failed for {}", src.display());
}

fn main() {
    if env::var_os("CARGO_FEATURE_HIP").is_none() {
        return;
    }
    let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").unwrap());
    let root = env::var("GSC_AMD_SRC")
        .map(PathBuf::from)
        .unwrap_or_else(|_| manifest.join("..").join("gstools-core_amd"));
    let csrc = root.join("csrc");
    let include = root.join("..").join("include");
    let out = PathBuf::from(env::var("OUT_DIR").unwrap());
    let mut objs = Vec::new();
    let mut srcs: Vec<PathBuf> = std::fs::read_dir(&csrc)
        .expect("reading csrc")
        .map(|e| e.unwrap().path())
        .filter(|p| p.extension().map_or(false, |x| x == "hip"))
        .collect();
    srcs.sort();
    for src in &srcs {
        let obj = out.join(src.file_stem().unwrap()).with_extension("o");
        compile(src, &include, &csrc, &obj);
        objs.push(obj);
        println!("cargo:rerun-if-changed={}", src.display());
    }
    let lib = out.join("libgstools_core_amd.so");
    let st = Command::new(hipcc())
        .args(["-shared", "-fPIC", "--offload-arch=gfx950", "-o"])
        .arg(&lib)
        .args(&objs)
        .status()
        .expect("linking with hipcc");
    assert!(st.success(), "link failed");
    println!("cargo:rustc-link-search=native={}", out.display());
    println!("cargo:rustc-link-lib=dylib=gstools_core_amd");
    // the Python extension module finds the library next to itself at run time
    println!("cargo:rustc-link-arg=-Wl,-rpath,$ORIGIN");
    println!("cargo:rerun-if-changed={}", include.join("gstools_core_amd.h").display());
    println!("cargo:rerun-if-env-changed=GSC_AMD_SRC");
    println!("cargo:rerun-if-env-changed=HIPCC");
}
