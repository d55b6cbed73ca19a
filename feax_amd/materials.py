"""Constitutive laws and surface loads the HIP element kernels implement.

In feax a Problem subclass returns a JAX function from get_tensor_map()
(feax/problem.py:595-627) or get_energy_density() (629-661) and the
element kernel differentiates it with jacfwd.  Here the kernels are
hand-written, so those hooks return one of the descriptors below.  Each is
still a plain callable with the same signature as the feax function
(``stress(u_grad, *params)`` / ``surface_map(u, x, *params)``), evaluated
with numpy, so user code reads the same; the device path reads its
parameters.
"""

import numpy as np

from . import _native as nat


class TensorMap:
    """Base class: a stress law sigma(u_grad, param) known to the kernels."""
    kind = None

    def material(self, param_loc):
        raise NotImplementedError


class IsotropicLinearElastic(TensorMap):
    """sigma = lam tr(eps) I + 2 mu eps with lam, mu from (E, nu).

    ``E`` fixed here, or (E=None) taken from TracedParams.volume_vars[0] as a
    node var (N,) or cell var (E,) -- the tensor map of
    examples/benchmark/bench_linear_elasticity.py:327-339 and
    tests/conftest.py:70-76.
    """
    kind = nat.FX_MAT_LINEAR_ELASTIC

    def __init__(self, nu=0.3, E=None):
        self.nu = float(nu)
        self.E = None if E is None else float(E)

    def __call__(self, u_grad, E=None):
        E = self.E if E is None else E
        lam = E * self.nu / ((1 + self.nu) * (1 - 2 * self.nu))
        mu = E / (2 * (1 + self.nu))
        eps = 0.5 * (u_grad + np.swapaxes(u_grad, -1, -2))
        d = np.shape(u_grad)[-1]
        return lam * np.trace(eps, axis1=-2, axis2=-1)[..., None, None] * np.eye(d) + 2 * mu * eps

    def material(self, param_loc):
        return nat.material_struct(self.kind, param_loc, nu=self.nu)


class SIMPLinearElastic(IsotropicLinearElastic):
    """Linear elastic with E(rho) = Emin + rho^penal (E0 - Emin), rho a cell
    var: the SIMP law of feax's topology-optimisation examples
    (feax/solvers/cmg.py:312-323 uses the same law)."""
    kind = nat.FX_MAT_SIMP

    def __init__(self, nu=0.3, E0=1.0, Emin=1e-9, penal=3.0):
        super().__init__(nu)
        self.E0, self.Emin, self.penal = float(E0), float(Emin), float(penal)

    def modulus(self, rho):
        return self.Emin + np.asarray(rho) ** self.penal * (self.E0 - self.Emin)

    def __call__(self, u_grad, rho):
        return IsotropicLinearElastic.__call__(self, u_grad, self.modulus(rho))

    def material(self, param_loc):
        return nat.material_struct(self.kind, param_loc, nu=self.nu, E0=self.E0, Emin=self.Emin,
                                   penal=self.penal)


class CompressibleNeoHookean(TensorMap):
    """P = mu (F - F^-T) + kappa (J - 1) J F^-T, psi = mu/2 (I1 - 3) - mu ln J +
    kappa/2 (J - 1)^2, mu a cell var (tests/test_newton_solver.py:65-89)."""
    kind = nat.FX_MAT_NEO_HOOKEAN

    def __init__(self, kappa=10.0):
        self.kappa = float(kappa)

    def __call__(self, u_grad, mu):
        F = np.eye(np.shape(u_grad)[-1]) + u_grad
        J = np.linalg.det(F)
        FiT = np.swapaxes(np.linalg.inv(F), -1, -2)
        return mu * (F - FiT) + self.kappa * (J - 1) * J * FiT

    def material(self, param_loc):
        return nat.material_struct(self.kind, param_loc, kappa=self.kappa)


class SurfaceTraction:
    """Dead traction t = direction * s with s the surface var at each face
    Gauss point: ``lambda u, x, t: np.array([0., 0., t])`` of
    tests/conftest.py:78 and the benchmark.  Added to the residual with a +
    sign (feax/assembler.py:529-533)."""

    def __init__(self, direction=(0.0, 0.0, 1.0)):
        self.direction = np.asarray(direction, float).reshape(-1)
        if self.direction.shape not in ((1,), (2,), (3,)):
            raise ValueError(f"traction direction must have 1, 2 or 3 components, got {self.direction}")

    def __call__(self, u, x, t):
        return self.direction * t


class IsotropicConduction(TensorMap):
    """Scalar flux q = k grad T (vec = 1): the tensor map
    ``lambda grad_T, k: k * grad_T`` of feax's tests/test_params_bridge.py:
    131-133 and examples/basic/thermo_mech_coupling.py:64-69.  ``k`` fixed
    here, or (k=None) taken from TracedParams.volume_vars[0] as a node or
    cell variable."""
    kind = nat.FX_MAT_CONDUCTION

    def __init__(self, k=None):
        self.k = None if k is None else float(k)

    @property
    def fixed_param(self):
        return self.k

    def __call__(self, grad_T, k=None):
        return (self.k if k is None else k) * grad_T

    def material(self, param_loc):
        return nat.material_struct(self.kind, param_loc)


class SIMPConduction(IsotropicConduction):
    """q = k(rho) grad T with k = kmin + rho^penal (k0 - kmin), rho a cell or
    node variable (thermal topology optimisation)."""
    kind = nat.FX_MAT_SIMP_CONDUCTION

    def __init__(self, k0=1.0, kmin=1e-3, penal=3.0):
        super().__init__(None)
        self.k0, self.kmin, self.penal = float(k0), float(kmin), float(penal)

    def conductivity(self, rho):
        return self.kmin + np.asarray(rho) ** self.penal * (self.k0 - self.kmin)

    def __call__(self, grad_T, rho):
        return self.conductivity(rho) * grad_T

    def material(self, param_loc):
        return nat.material_struct(self.kind, param_loc, E0=self.k0, Emin=self.kmin, penal=self.penal)


class SurfaceFlux(SurfaceTraction):
    """Prescribed normal flux q = scale * s on a loaded boundary of a scalar
    problem (the one-component surface map ``lambda u, x, q: np.array([q])``),
    added with a + sign like a traction."""

    def __init__(self, scale=1.0):
        super().__init__((float(scale),))


class ThermoElastic(IsotropicLinearElastic):
    """sigma = lam tr(eps_m) I + 2 mu eps_m with eps_m = eps - alpha (T - t_ref) I:
    the thermo-elastic tensor map ``stress(u_grad, E, T)`` of feax's
    tests/test_params_bridge.py:136-143 and examples/basic/
    thermo_mech_coupling.py:95-103.  E is volume_vars[0] (node or cell
    variable, or fixed with E=...) and the temperature T is volume_vars[1],
    a node variable (e.g. TracedParams.node_var_from_solution of a thermal
    solve).  HEX8 only."""

    def __init__(self, nu=0.3, alpha=1e-5, t_ref=0.0, E=None):
        super().__init__(nu, E)
        self.alpha, self.t_ref = float(alpha), float(t_ref)

    def __call__(self, u_grad, E=None, T=0.0):
        u_grad = np.asarray(u_grad, float)
        shift = self.alpha * (np.asarray(T, float) - self.t_ref)
        return super().__call__(u_grad - shift[..., None, None] * np.eye(u_grad.shape[-1]), E)


class SIMPThermoElastic(SIMPLinearElastic):
    """ThermoElastic with E = Emin + rho^penal (E0 - Emin) (the lattice law of
    examples/basic/thermo_mech_coupling.py:95-103 with Emin = 0)."""

    def __init__(self, nu=0.3, E0=1.0, Emin=1e-9, penal=3.0, alpha=1e-5, t_ref=0.0):
        super().__init__(nu, E0, Emin, penal)
        self.alpha, self.t_ref = float(alpha), float(t_ref)

    def __call__(self, u_grad, rho, T=0.0):
        E = self.modulus(rho)
        return ThermoElastic(self.nu, self.alpha, self.t_ref)(u_grad, E, T)


class ThermoElasticCoupled(TensorMap):
    """The weak form of a monolithic two-variable (u, T) problem
    (Problem(mesh=[m, m], vec=[3, 1]); feax's get_weak_form,
    feax/problem.py:697-740, feax/assembler.py:591-647), evaluated by the
    fx_hex8ut kernels:
        grad terms [sigma, k grad T], mass terms [0, 0],
        sigma = E (lam0 tr eps I + 2 mu0 eps) - E alpha (T - t_ref) / (1 - 2 nu) I,
    E = volume_vars[0] (node or cell variable, or fixed with E=...),
    optionally through the SIMP law (simp=(E0, Emin, penal)); nu, alpha,
    t_ref and the conductivity k are constants.  Return it from a Problem's
    get_weak_form()."""

    kind = nat.FX_MAT_LINEAR_ELASTIC
    coupled = True

    def __init__(self, nu=0.3, alpha=1e-5, t_ref=0.0, k=1.0, E=None, simp=None):
        self.nu, self.alpha, self.t_ref, self.k = float(nu), float(alpha), float(t_ref), float(k)
        self.E = E
        self.simp = simp

    def weak_form(self, vals, grads, x, E):
        """(mass_terms, grad_terms) at one point: the function feax would vmap."""
        T = float(np.asarray(vals[1]).reshape(-1)[0])
        ug = np.asarray(grads[0], float)
        if self.simp is not None:
            E0, Emin, penal = self.simp
            E = Emin + E ** penal * (E0 - Emin)
        eps = 0.5 * (ug + ug.T)
        lam0 = self.nu / ((1 + self.nu) * (1 - 2 * self.nu))
        mu0 = 1.0 / (2 * (1 + self.nu))
        sig = E * (lam0 * np.trace(eps) * np.eye(3) + 2 * mu0 * eps)
        sig -= E * self.alpha * (T - self.t_ref) / (1 - 2 * self.nu) * np.eye(3)
        return [np.zeros(3), np.zeros(1)], [sig, self.k * np.asarray(grads[1], float).reshape(1, 3)]

    def material(self, param_loc):
        if self.simp is not None:
            E0, Emin, penal = self.simp
            return nat.material_struct(nat.FX_MAT_SIMP, param_loc, nu=self.nu, E0=E0, Emin=Emin, penal=penal,
                                       kappa=self.k, alpha=self.alpha, t_ref=self.t_ref)
        return nat.material_struct(self.kind, param_loc, nu=self.nu, kappa=self.k, alpha=self.alpha,
                                   t_ref=self.t_ref)
