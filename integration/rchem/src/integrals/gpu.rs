//! The MI355X library (librchem_gpu.so, include/rchem_gpu.h of this
//! repository) behind rchem's integral entry points, compiled with the `gpu`
//! feature.  Included from src/integrals/mod.rs as `pub mod gpu;`, next to
//! `tho66` (tho66.rs:225-260 binds libpyquante2 the same way).
//!
//! Written for the rchem crate; not compiled in the image this repository was
//! built in (no cargo there).
#![allow(non_camel_case_types, non_snake_case, non_upper_case_globals, dead_code)]

pub mod ffi {
    include!(concat!(env!("OUT_DIR"), "/bindings_rchem_gpu.rs"));
}

use ndarray::{Array, Ix2, Ix4};
use std::ffi::CStr;

#[derive(Debug)]
pub struct GpuError {
    pub code: i32,
    pub message: String,
}

/// One shell as the library takes it: centre, angular momentum, exponents and
/// the raw contraction coefficients (the library applies rchem's per-component
/// primitive norms, basis.rs:140-149).
#[derive(Clone, Debug)]
pub struct Shell {
    pub origin: [f64; 3],
    pub l: i32,
    pub exponents: Vec<f64>,
    pub coefs: Vec<f64>,
}

#[derive(Clone, Copy, Debug)]
pub struct ScfResult {
    pub e_elec: f64,
    pub e_nuc: f64,
    pub e_total: f64,
    pub iterations: i32,
    pub converged: bool,
}

/// A device context (one HIP stream on one GPU) holding the current basis.
pub struct Gpu {
    ctx: *mut ffi::rchem_ctx,
    nbf: usize,
}

fn check(ctx: *const ffi::rchem_ctx, rc: i32) -> Result<(), GpuError> {
    if rc == ffi::rchem_status_RCHEM_OK as i32 {
        return Ok(());
    }
    let message = unsafe { CStr::from_ptr(ffi::rchem_last_error(ctx)) }
        .to_string_lossy()
        .into_owned();
    Err(GpuError { code: rc, message })
}

fn contiguous(a: &Array<f64, Ix2>) -> bool {
    a.is_standard_layout()
}

impl Gpu {
    pub fn new(device: i32) -> Result<Gpu, GpuError> {
        let mut ctx = std::ptr::null_mut();
        check(std::ptr::null(), unsafe { ffi::rchem_ctx_create(device, &mut ctx) })?;
        Ok(Gpu { ctx, nbf: 0 })
    }

    pub fn nbf(&self) -> usize {
        self.nbf
    }

    /// rchem_basis_upload: shells in rchem's function order (atom, BSE shell,
    /// l, get_ijk_list component; basis.rs:186-205).
    pub fn upload_shells(&mut self, shells: &[Shell]) -> Result<(), GpuError> {
        let mut xyz = Vec::with_capacity(3 * shells.len());
        let mut l = Vec::with_capacity(shells.len());
        let mut nprim = Vec::with_capacity(shells.len());
        let (mut exps, mut coefs) = (Vec::new(), Vec::new());
        for s in shells {
            xyz.extend_from_slice(&s.origin);
            l.push(s.l);
            nprim.push(s.exponents.len() as i32);
            exps.extend_from_slice(&s.exponents);
            coefs.extend_from_slice(&s.coefs);
        }
        check(self.ctx, unsafe {
            ffi::rchem_basis_upload(
                self.ctx,
                shells.len() as i32,
                xyz.as_ptr(),
                l.as_ptr(),
                nprim.as_ptr(),
                exps.as_ptr(),
                coefs.as_ptr(),
            )
        })?;
        let mut nbf = 0i32;
        check(self.ctx, unsafe { ffi::rchem_basis_nbf(self.ctx, &mut nbf) })?;
        self.nbf = nbf as usize;
        Ok(())
    }

    /// JK_direct semantics (basis.rs:383-428): J and K overwritten.
    pub fn jk(
        &mut self,
        D: &Array<f64, Ix2>,
        J: &mut Array<f64, Ix2>,
        K: &mut Array<f64, Ix2>,
    ) -> Result<(), GpuError> {
        assert!(contiguous(D) && contiguous(J) && contiguous(K));
        assert!(D.dim() == (self.nbf, self.nbf) && J.dim() == D.dim() && K.dim() == D.dim());
        check(self.ctx, unsafe {
            ffi::rchem_jk(self.ctx, D.as_ptr(), J.as_mut_ptr(), K.as_mut_ptr(), std::ptr::null())
        })
    }

    /// S, T and V (basis.rs:234, :273, :316).
    pub fn one_electron(
        &mut self,
        atomcoords: &[[f64; 3]],
        atomnos: &[u64],
    ) -> Result<(Array<f64, Ix2>, Array<f64, Ix2>, Array<f64, Ix2>), GpuError> {
        let n = self.nbf;
        let mut S = Array::zeros((n, n));
        let mut T = Array::zeros((n, n));
        let mut V = Array::zeros((n, n));
        let xyz: Vec<f64> = atomcoords.iter().flatten().copied().collect();
        let Z: Vec<f64> = atomnos.iter().map(|&z| z as f64).collect();
        check(self.ctx, unsafe {
            ffi::rchem_one_electron(
                self.ctx,
                Z.len() as i32,
                Z.as_ptr(),
                xyz.as_ptr(),
                S.as_mut_ptr(),
                T.as_mut_ptr(),
                V.as_mut_ptr(),
            )
        })?;
        Ok((S, T, V))
    }

    /// build_I (basis.rs:430-460): the n^4 tensor.
    pub fn eri_tensor(&mut self) -> Result<Array<f64, Ix4>, GpuError> {
        let n = self.nbf;
        let mut I = Array::zeros((n, n, n, n));
        check(self.ctx, unsafe {
            ffi::rchem_eri_tensor(self.ctx, I.as_mut_ptr(), std::ptr::null())
        })?;
        Ok(I)
    }

    fn scf_opts(&self, nocc: usize) -> ffi::rchem_scf_opts {
        let mut o: ffi::rchem_scf_opts = unsafe { std::mem::zeroed() };
        unsafe { ffi::rchem_scf_opts_default(&mut o) };
        o.nocc = nocc as i32;
        o
    }

    /// rchem's SCF (rchem.rs:40-90) on the device.
    pub fn scf(
        &mut self,
        atomcoords: &[[f64; 3]],
        atomnos: &[u64],
        nocc: usize,
    ) -> Result<(ScfResult, Array<f64, Ix2>), GpuError> {
        let o = self.scf_opts(nocc);
        let xyz: Vec<f64> = atomcoords.iter().flatten().copied().collect();
        let Z: Vec<f64> = atomnos.iter().map(|&z| z as f64).collect();
        let mut r: ffi::rchem_scf_result = unsafe { std::mem::zeroed() };
        let mut D = Array::zeros((self.nbf, self.nbf));
        check(self.ctx, unsafe {
            ffi::rchem_scf(self.ctx, Z.len() as i32, Z.as_ptr(), xyz.as_ptr(), &o, &mut r,
                           D.as_mut_ptr())
        })?;
        Ok((to_result(&r), D))
    }

    /// dE_total/dR (Eh/bohr), natom x 3, after the SCF it runs itself.
    pub fn rhf_gradient(
        &mut self,
        atomcoords: &[[f64; 3]],
        atomnos: &[u64],
        nocc: usize,
    ) -> Result<(ScfResult, Array<f64, Ix2>), GpuError> {
        let o = self.scf_opts(nocc);
        let xyz: Vec<f64> = atomcoords.iter().flatten().copied().collect();
        let Z: Vec<f64> = atomnos.iter().map(|&z| z as f64).collect();
        let mut r: ffi::rchem_scf_result = unsafe { std::mem::zeroed() };
        let mut g = Array::zeros((Z.len(), 3));
        check(self.ctx, unsafe {
            ffi::rchem_rhf_gradient(self.ctx, Z.len() as i32, Z.as_ptr(), xyz.as_ptr(), &o,
                                    &mut r, g.as_mut_ptr())
        })?;
        Ok((to_result(&r), g))
    }
}

fn to_result(r: &ffi::rchem_scf_result) -> ScfResult {
    ScfResult {
        e_elec: r.e_elec,
        e_nuc: r.e_nuc,
        e_total: r.e_total,
        iterations: r.iterations,
        converged: r.converged != 0,
    }
}

impl Drop for Gpu {
    fn drop(&mut self) {
        unsafe {
            ffi::rchem_ctx_destroy(self.ctx);
        }
    }
}
