//! `hip` module of the reference crate: the summators and kriging calculators
//! on one MI355X through the C-ABI of libgstools_core_amd.so
//! (include/gstools_core_amd.h), with the signatures of src/field.rs and
//! src/krige.rs.  Drop into the crate's src/, declare `#[cfg(feature = "hip")]
//! mod hip;` in src/lib.rs and route the PyO3 entry points (lib.rs:47, 64, 83,
//! 99, 117) to these functions; build.rs compiles and links the library.
//!
//! ndarray views pass as a pointer and signed element strides, which is what
//! the C-ABI takes, so no view is copied on the Rust side.  A failing call
//! panics with the library's message, as the reference panics on bad shapes.
use ndarray::{Array1, Array2, ArrayView1, ArrayView2, ShapeBuilder};
use std::ffi::CStr;
use std::os::raw::{c_char, c_int, c_void};

/// gsc_opts (include/gstools_core_amd.h): filled by gsc_opts_init.
#[repr(C)]
pub struct GscOpts {
    pub out_dtype: i32,
    pub path: i32,
    pub precision: i32,
    pub device: i32,
    pub num_threads: i32,
    pub accum: i32,
    pub direct_variant: i32,
    pub blocks_per_cu: i32,
    pub scale: f64,
    pub tile: i32,
    pub table_bytes: i64,
}

extern "C" {
    fn gsc_opts_init(opts: *mut GscOpts);
    fn gsc_last_error() -> *const c_char;
    fn gsc_summate(dim: c_int, m: i64, n: i64,
                   cov: *const f64, cov_sd: i64, cov_sm: i64,
                   z1: *const f64, z1_s: i64, z2: *const f64, z2_s: i64,
                   pos: *const f64, pos_sd: i64, pos_sp: i64,
                   out: *mut c_void, out_s: i64, opts: *const GscOpts) -> c_int;
    fn gsc_summate_incompr(dim: c_int, m: i64, n: i64,
                           cov: *const f64, cov_sd: i64, cov_sm: i64,
                           z1: *const f64, z1_s: i64, z2: *const f64, z2_s: i64,
                           pos: *const f64, pos_sd: i64, pos_sp: i64,
                           out: *mut c_void, out_sd: i64, out_sp: i64,
                           opts: *const GscOpts) -> c_int;
    fn gsc_summate_fourier(dim: c_int, m: i64, n: i64, sf: *const f64, sf_s: i64,
                           modes: *const f64, md_sd: i64, md_sm: i64,
                           z1: *const f64, z1_s: i64, z2: *const f64, z2_s: i64,
                           pos: *const f64, pos_sd: i64, pos_sp: i64,
                           out: *mut c_void, out_s: i64, opts: *const GscOpts) -> c_int;
    fn gsc_krige(n: i64, nt: i64, mat: *const f64, mat_s0: i64, mat_s1: i64,
                 vecs: *const f64, vecs_s0: i64, vecs_s1: i64, cond: *const f64, cond_s: i64,
                 field: *mut f64, field_s: i64, error: *mut f64, error_s: i64,
                 opts: *const GscOpts) -> c_int;
}

fn check(rc: c_int) {
    if rc != 0 {
        let msg = unsafe { CStr::from_ptr(gsc_last_error()) }.to_string_lossy();
        panic!("gstools_core_amd: {msg}");
    }
}

fn opts(num_threads: Option<usize>) -> GscOpts {
    let mut o = std::mem::MaybeUninit::<GscOpts>::uninit();
    let mut opts = unsafe {
        gsc_opts_init(o.as_mut_ptr());
        o.assume_init()
    };
    // the host-side staging threads; the device does the summation
    opts.num_threads = num_threads.unwrap_or(0) as i32;
    opts
}

/// Same signature and result as field::summator (src/field.rs:37-43).
pub fn summator(cov_samples: ArrayView2<'_, f64>, z1: ArrayView1<'_, f64>,
                z2: ArrayView1<'_, f64>, pos: ArrayView2<'_, f64>,
                num_threads: Option<usize>) -> Array1<f64> {
    assert_eq!(cov_samples.dim().0, pos.dim().0);
    assert_eq!(cov_samples.dim().1, z1.dim());
    assert_eq!(cov_samples.dim().1, z2.dim());
    let (d, m) = cov_samples.dim();
    let n = pos.dim().1;
    let mut out = Array1::<f64>::zeros(n);
    let o = opts(num_threads);
    let (cs, s1, s2, ps) = (cov_samples.strides(), z1.strides(), z2.strides(), pos.strides());
    check(unsafe {
        gsc_summate(d as c_int, m as i64, n as i64,
                    cov_samples.as_ptr(), cs[0] as i64, cs[1] as i64,
                    z1.as_ptr(), s1[0] as i64, z2.as_ptr(), s2[0] as i64,
                    pos.as_ptr(), ps[0] as i64, ps[1] as i64,
                    out.as_mut_ptr() as *mut c_void, 1, &o)
    });
    out
}

/// Same signature and result as field::summator_incompr (src/field.rs:97-103):
/// (d, N) in Fortran order (field.rs:166-174).
pub fn summator_incompr(cov_samples: ArrayView2<'_, f64>, z1: ArrayView1<'_, f64>,
                        z2: ArrayView1<'_, f64>, pos: ArrayView2<'_, f64>,
                        num_threads: Option<usize>) -> Array2<f64> {
    assert_eq!(cov_samples.dim().0, pos.dim().0);
    assert_eq!(cov_samples.dim().1, z1.dim());
    assert_eq!(cov_samples.dim().1, z2.dim());
    let (d, m) = cov_samples.dim();
    let n = pos.dim().1;
    let mut out = Array2::<f64>::zeros((d, n).f());
    let o = opts(num_threads);
    let (cs, s1, s2, ps) = (cov_samples.strides(), z1.strides(), z2.strides(), pos.strides());
    let os = out.strides().to_vec();
    check(unsafe {
        gsc_summate_incompr(d as c_int, m as i64, n as i64,
                            cov_samples.as_ptr(), cs[0] as i64, cs[1] as i64,
                            z1.as_ptr(), s1[0] as i64, z2.as_ptr(), s2[0] as i64,
                            pos.as_ptr(), ps[0] as i64, ps[1] as i64,
                            out.as_mut_ptr() as *mut c_void, os[0] as i64, os[1] as i64, &o)
    });
    out
}

/// Same signature and result as field::summator_fourier (src/field.rs:219-226).
pub fn summator_fourier(spectrum_factor: ArrayView1<'_, f64>, modes: ArrayView2<'_, f64>,
                        z1: ArrayView1<'_, f64>, z2: ArrayView1<'_, f64>,
                        pos: ArrayView2<'_, f64>, num_threads: Option<usize>) -> Array1<f64> {
    assert_eq!(modes.dim().0, pos.dim().0);
    assert_eq!(modes.dim().1, z1.dim());
    assert_eq!(modes.dim().1, z2.dim());
    let (d, m) = modes.dim();
    let n = pos.dim().1;
    let mut out = Array1::<f64>::zeros(n);
    let o = opts(num_threads);
    let (fs, ms, s1, s2, ps) =
        (spectrum_factor.strides(), modes.strides(), z1.strides(), z2.strides(), pos.strides());
    check(unsafe {
        gsc_summate_fourier(d as c_int, m as i64, n as i64,
                            spectrum_factor.as_ptr(), fs[0] as i64,
                            modes.as_ptr(), ms[0] as i64, ms[1] as i64,
                            z1.as_ptr(), s1[0] as i64, z2.as_ptr(), s2[0] as i64,
                            pos.as_ptr(), ps[0] as i64, ps[1] as i64,
                            out.as_mut_ptr() as *mut c_void, 1, &o)
    });
    out
}

fn krige(krig_mat: ArrayView2<'_, f64>, krig_vecs: ArrayView2<'_, f64>,
         cond: ArrayView1<'_, f64>, num_threads: Option<usize>,
         with_error: bool) -> (Array1<f64>, Array1<f64>) {
    assert_eq!(krig_mat.dim().0, krig_mat.dim().1);
    assert_eq!(krig_mat.dim().0, krig_vecs.dim().0);
    assert_eq!(krig_mat.dim().0, cond.dim());
    let (n, nt) = krig_vecs.dim();
    let mut field = Array1::<f64>::zeros(nt);
    let mut error = Array1::<f64>::zeros(if with_error { nt } else { 0 });
    let o = opts(num_threads);
    let (ms, vs, cs) = (krig_mat.strides(), krig_vecs.strides(), cond.strides());
    let err_ptr = if with_error { error.as_mut_ptr() } else { std::ptr::null_mut() };
    check(unsafe {
        gsc_krige(n as i64, nt as i64, krig_mat.as_ptr(), ms[0] as i64, ms[1] as i64,
                  krig_vecs.as_ptr(), vs[0] as i64, vs[1] as i64, cond.as_ptr(), cs[0] as i64,
                  field.as_mut_ptr(), 1, err_ptr, 1, &o)
    });
    (field, error)
}

/// Same signature and result as krige::calculator_field_krige_and_variance
/// (src/krige.rs:24-29).
pub fn calculator_field_krige_and_variance(krig_mat: ArrayView2<'_, f64>,
                                           krig_vecs: ArrayView2<'_, f64>,
                                           cond: ArrayView1<'_, f64>,
                                           num_threads: Option<usize>)
                                           -> (Array1<f64>, Array1<f64>) {
    krige(krig_mat, krig_vecs, cond, num_threads, true)
}

/// Same signature and result as krige::calculator_field_krige (src/krige.rs:92-97).
pub fn calculator_field_krige(krig_mat: ArrayView2<'_, f64>, krig_vecs: ArrayView2<'_, f64>,
                              cond: ArrayView1<'_, f64>, num_threads: Option<usize>)
                              -> Array1<f64> {
    krige(krig_mat, krig_vecs, cond, num_threads, false).0
}
