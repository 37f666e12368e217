//! GPU bodies of rchem's basis-level entry points (basis.rs:234-484), for the
//! `gpu` feature.  Basis's fields are private (basis.rs:176-179), so this file
//! is appended to src/basis.rs (it uses CGTO and Basis directly); with the
//! feature on, the reference's own `S`, `T`, `V`, `JK_direct` and `build_I`
//! are compiled out (`#[cfg(not(feature = "gpu"))]` on each) and these take
//! their place with the same signatures, so src/bin/rchem.rs is unchanged.
//!
//! Written for the rchem crate; not compiled in the image this repository was
//! built in (no cargo there).

#[cfg(feature = "gpu")]
use crate::integrals::gpu::{Gpu, Shell};
#[cfg(feature = "gpu")]
use std::cell::RefCell;

#[cfg(feature = "gpu")]
impl Basis {
    /// The shells behind the Cartesian functions: Basis::new pushes, per BSE
    /// shell and l, the get_ijk_list(l) components in order (basis.rs:189-205),
    /// so a shell of angular momentum l is the run of (l+1)(l+2)/2 functions
    /// starting at a component [l, 0, 0].
    pub(crate) fn gpu_shells(&self) -> Vec<Shell> {
        let mut out = Vec::new();
        let mut i = 0;
        while i < self.cgtos.len() {
            let c = &self.cgtos[i];
            let l: usize = c.powers.iter().sum();
            assert_eq!(c.powers, [l, 0, 0], "function {} does not start a shell", i);
            out.push(Shell {
                origin: c.origin,
                l: l as i32,
                exponents: c.primitives.iter().map(|p| p.exponent).collect(),
                coefs: c.coefs.clone(),
            });
            i += (l + 1) * (l + 2) / 2;
        }
        out
    }
}

#[cfg(feature = "gpu")]
thread_local! {
    // one device context per thread, holding the basis last uploaded
    static GPU: RefCell<Option<(usize, Gpu)>> = RefCell::new(None);
}

/// Runs f on this thread's context with `basis_set` uploaded (re-uploaded only
/// when another Basis is passed).
#[cfg(feature = "gpu")]
fn with_gpu<R>(basis_set: &Basis, f: impl FnOnce(&mut Gpu) -> R) -> R {
    GPU.with(|cell| {
        let mut slot = cell.borrow_mut();
        let key = basis_set as *const Basis as usize;
        if slot.as_ref().map(|(k, _)| *k) != Some(key) {
            let mut g = match slot.take() {
                Some((_, g)) => g,
                None => Gpu::new(0).expect("rchem_ctx_create"),
            };
            g.upload_shells(&basis_set.gpu_shells()).expect("rchem_basis_upload");
            *slot = Some((key, g));
        }
        f(&mut slot.as_mut().unwrap().1)
    })
}

#[cfg(feature = "gpu")]
pub fn S(basis_set: &Basis) -> Array<f64, Ix2> {
    with_gpu(basis_set, |g| g.one_electron(&[], &[]).expect("rchem_one_electron").0)
}

#[cfg(feature = "gpu")]
pub fn T(basis_set: &Basis) -> Array<f64, Ix2> {
    with_gpu(basis_set, |g| g.one_electron(&[], &[]).expect("rchem_one_electron").1)
}

#[cfg(feature = "gpu")]
pub fn V(basis_set: &Basis, atomcoords: &[[f64; 3]], atomnos: &Vec<u64>) -> Array<f64, Ix2> {
    with_gpu(basis_set, |g| g.one_electron(atomcoords, atomnos).expect("rchem_one_electron").2)
}

#[cfg(feature = "gpu")]
pub fn JK_direct(
    J: &mut Array<f64, Ix2>,
    K: &mut Array<f64, Ix2>,
    basis_set: &Basis,
    D: &Array<f64, Ix2>,
) {
    // D may be a view with another layout (rchem builds it with .dot, which
    // gives standard layout); J and K are the caller's, overwritten
    let Dc = D.as_standard_layout().to_owned();
    with_gpu(basis_set, |g| g.jk(&Dc, J, K).expect("rchem_jk"))
}

#[cfg(feature = "gpu")]
pub fn build_I(basis_set: &Basis) -> Array<f64, Ix4> {
    with_gpu(basis_set, |g| g.eri_tensor().expect("rchem_eri_tensor"))
}

/// rchem's RHF energy as an operation with a derivative: forward() is
/// E_elec + E_nuc after the SCF of rchem.rs:40-90; backward() is dE/dR
/// (Eh/bohr, natom x 3) from the adjoint kernels of rchem_rhf_gradient, which
/// contract derivative integrals with D and W = C_occ eps C_occ^T, recomputing
/// every quartet instead of storing it.
#[cfg(feature = "gpu")]
pub struct RhfEnergy<'a> {
    pub basis_set: &'a Basis,
    pub atomcoords: Vec<[f64; 3]>,
    pub atomnos: Vec<u64>,
    pub nocc: usize,
}

#[cfg(feature = "gpu")]
impl<'a> RhfEnergy<'a> {
    pub fn forward(&self) -> f64 {
        with_gpu(self.basis_set, |g| {
            let (r, _) = g.scf(&self.atomcoords, &self.atomnos, self.nocc).expect("rchem_scf");
            assert!(r.converged, "SCF did not converge");
            r.e_total
        })
    }

    pub fn backward(&self) -> Array<f64, Ix2> {
        with_gpu(self.basis_set, |g| {
            let (r, grad) = g
                .rhf_gradient(&self.atomcoords, &self.atomnos, self.nocc)
                .expect("rchem_rhf_gradient");
            assert!(r.converged, "SCF did not converge");
            grad
        })
    }
}
