"""CPU restatement of the time-loop steps either side of the Stokes solve
(test infrastructure only): the material law of mfiles_M3TET/rock_prop.m:31-118
and the advective time step of mfiles_M3TET/calculate_dt.m:40-56."""
import numpy as np

R_GAS = 8.314472  # rock_prop.m:68


def rock_prop(T, X, Visc0, Dens0, T_ref, alpha=0.0, Ea=0.0, Va=0.0, g=-1.0, L0=1.0, a_UM2LM=0.0,
              minVisc=None, maxVisc=None, minDens=None, maxDens=None):
    """Nodal (Visc, Dens) from nodal T; X are the node coordinates (n x 3)."""
    T = np.asarray(T, dtype=float)
    radius = np.linalg.norm(X, axis=1)                 # rock_prop.m:36-37
    depth = radius.max() - radius
    dens = Dens0 * (1 - alpha * (T - T_ref))           # :46
    if maxDens is not None:
        dens = np.minimum(dens, maxDens)               # :56
    if minDens is not None:
        dens = np.maximum(dens, minDens)               # :59
    v = np.ones_like(T)
    T0K = T_ref + 273
    if Ea > 0:                                         # :71-85
        if Ea < 1000:
            raise ValueError("Activation energy PHYSICS.Ea seems to be in kJ/mol. Convert to J/mol.")
        v = v * np.exp(Ea / (R_GAS * T0K) * (T0K / (T + 273) - 1))
        v = np.maximum(np.minimum(v, 1e8), 1e-8)
    v2 = v.copy()
    if Va > 0:                                         # :89-96
        P = np.abs(g * Dens0 * L0 * depth)
        v2 = v2 * np.exp(P * Va / (R_GAS * T0K))
    v3 = v.copy()
    if a_UM2LM > 0:                                    # :100-103
        v3 = v3 * a_UM2LM
    visc = np.minimum(v2, v3) * Visc0                  # :107-111
    if maxVisc is not None:
        visc = np.minimum(visc, maxVisc)               # :117
    if minVisc is not None:
        visc = np.maximum(visc, minVisc)               # :120
    return visc, dens


def calculate_dt(Uxyz, X, dt_adv_limit, dtmax):
    """calculate_dt.m:40-56 with Uxyz (n x 3) Cartesian nodal velocities and
    L = mesh extents (init_mesh_curved_p.m:270)."""
    m = np.abs(Uxyz).max(axis=0)
    if np.all(m == 0):
        return dtmax
    L = X.max(axis=0) - X.min(axis=0)
    with np.errstate(divide="ignore"):
        dta = np.min(np.where(m > 0, L / np.where(m > 0, m, 1), np.inf))
    return min(dtmax, dt_adv_limit * dta)
