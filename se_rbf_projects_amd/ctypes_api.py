"""ctypes mirrors of the C types in include/serbf.h and include/serbf_ref.h and the
method/flow presets of the benchmark configurations."""
import ctypes as C

import numpy as np

c_double_p = C.POINTER(C.c_double)
c_int_p = C.POINTER(C.c_int)

MODE_REFERENCE, MODE_CONFIG = 0, 1
FLOW_TC2, FLOW_SOLID_BODY, FLOW_DEFORM = 0, 1, 2
F64, F32 = 0, 1
EARTH_R = 6.37122e6
DAY = 86400.0


class SerbfConfig(C.Structure):
    _fields_ = [
        ("Nv", C.c_int),
        ("htop", C.c_double),
        ("R", C.c_double),
        ("mode", C.c_int),
        ("n", C.c_int),
        ("phs_order", C.c_int),
        ("poly_deg", C.c_int),
        ("hv_order", C.c_int),
        ("hv_phs_order", C.c_int),
        ("hv_eps", C.c_double),
        ("nv", C.c_int),
        ("v_phs_order", C.c_int),
        ("v_poly_deg", C.c_int),
        ("hv_coef", C.c_double),
        ("flow", C.c_int),
        ("u0", C.c_double),
        ("alpha", C.c_double),
        ("T", C.c_double),
        ("k_def", C.c_double),
        ("w0", C.c_double),
        ("ntracers", C.c_int),
    ]


def make_config(**kw):
    cfg = SerbfConfig()
    defaults = dict(Nv=10, htop=12000.0, R=EARTH_R, mode=MODE_CONFIG, n=31, phs_order=7, poly_deg=3,
                    hv_order=4, hv_phs_order=9, hv_eps=0.0, nv=5, v_phs_order=5, v_poly_deg=3,
                    hv_coef=0.0, flow=FLOW_SOLID_BODY, u0=2 * np.pi * EARTH_R / (12 * DAY),
                    alpha=np.pi / 4, T=12 * DAY, k_def=0.0, w0=0.0, ntracers=1)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(cfg, k, v)
    return cfg


def reference_config(N, n=55, Nv=32, htop=12000.0):
    """Reference solver settings: PHS r^5, degree-5 tangent monomials, Gaussian
    Delta^4 with eps = 0.044 sqrt(N) - 0.14 (wrappers.c:209-211), TC2 flow."""
    return make_config(Nv=Nv, htop=htop, mode=MODE_REFERENCE, n=n, phs_order=5, poly_deg=5, hv_order=4,
                       hv_phs_order=0, hv_eps=0.044 * np.sqrt(N) - 0.14, nv=7, v_phs_order=0,
                       v_poly_deg=0, hv_coef=0.0, flow=FLOW_TC2, u0=40.0, alpha=0.0, T=86400.0,
                       k_def=5.0, w0=0.15)


def hex_spacing(N):
    """Mean spacing of N quasi-uniform nodes on the unit sphere (hexagonal cells)."""
    return float(np.sqrt(8 * np.pi / (np.sqrt(3) * N)))


# hyperviscosity coefficient gamma_c of hv_coef = -gamma_c h^8 u0 / (R h) (see hv_coefficient)
HV_GAMMA = 2.0 ** -14


def reference_eps(N):
    """Gaussian shape parameter of the unit sphere, eps = 0.044 sqrt(N) - 0.14 (wrappers.c:209-211)."""
    return 0.044 * float(np.sqrt(N)) - 0.14


def hv_coefficient(N, u0, gamma_c=None):
    """hv_coef = -gamma_c h^8 u0 / (R h): the dimensionless -gamma_c h^8 of the configs applied as
    a rate, in units of the cell-crossing frequency u0 / (R h)."""
    h = hex_spacing(N)
    g = HV_GAMMA if gamma_c is None else gamma_c
    return -g * h ** 8 * u0 / (EARTH_R * h)


def _hv_fields(N, hv):
    """hv = "gauss": Gaussian + constant Delta^4 with the reference's eps; "phs": PHS r^9 + the
    gradient monomials."""
    if hv == "gauss":
        return dict(hv_phs_order=0, hv_eps=reference_eps(N))
    if hv == "phs":
        return dict(hv_phs_order=9, hv_eps=0.0)
    raise ValueError(hv)


def config_A(N=2562, hv="gauss", gamma_c=None):
    """Plain method, CPU-runnable: ICOS level 4, Nv = 10 to 12 km, n = 31, PHS r^7 +
    degree-3 tangent monomials, vertical n = 5, Delta^4 with gamma = -2^-14 h^8,
    tilted (45 deg) solid-body rotation with u0 = 2 pi R / 12 days."""
    u0 = 2 * np.pi * EARTH_R / (12 * DAY)
    return make_config(Nv=10, htop=12000.0, mode=MODE_CONFIG, n=31, phs_order=7, poly_deg=3, hv_order=4,
                       nv=5, v_phs_order=5, v_poly_deg=3, hv_coef=hv_coefficient(N, u0, gamma_c),
                       flow=FLOW_SOLID_BODY, u0=u0, alpha=np.pi / 4, T=12 * DAY, **_hv_fields(N, hv))


def config_deform(N, Nv=60, n=55, deg=4, nv=7, hv="gauss", gamma_c=None):
    """Deformational flow of the at-scale configs: solid-body rotation 2 pi R / 12 days
    plus a reversing swirl of amplitude 10 R / T and w = 0.15 sin(pi z/ztop) cos(pi t/T)."""
    T = 12 * DAY
    u0 = 2 * np.pi * EARTH_R / T
    return make_config(Nv=Nv, htop=12000.0, mode=MODE_CONFIG, n=n, phs_order=7, poly_deg=deg, hv_order=4,
                       nv=nv, v_phs_order=5, v_poly_deg=nv - 2, hv_coef=hv_coefficient(N, u0, gamma_c),
                       flow=FLOW_DEFORM, u0=u0, alpha=0.0, T=T, k_def=10 * EARTH_R / T, w0=0.15,
                       **_hv_fields(N, hv))


def variable_stencil_sizes(D, n_min=31, n_max=84, k=10):
    """Per-node stencil size from local node density and the polynomial degree that goes with
    it.  Density is (1 / r_k)^2 with r_k the distance to the k-th neighbour (D: sorted stencil
    distances [N, >k]); sizes spread over [n_min, n_max] by the rank of the density (the
    densest nodes get the largest stencils).  Degree 3 below 45 nodes, 4 below 63, else 5
    (about three nodes per polynomial term)."""
    rk = np.asarray(D)[:, k]
    order = np.argsort(rk, kind="stable")[::-1]   # sparsest (largest r_k) first
    rank = np.empty(len(rk))
    rank[order] = np.arange(len(rk))
    frac = rank / max(len(rk) - 1, 1)
    nn = np.rint(n_min + (n_max - n_min) * frac).astype(np.int32)
    deg = np.where(nn < 45, 3, np.where(nn < 63, 4, 5)).astype(np.int32)
    return nn, deg


def courant_dt(cfg, N, courant=0.5):
    """dt = courant * h R / u_max with u_max the largest horizontal speed of the flow."""
    h = hex_spacing(N)
    if cfg.flow == FLOW_SOLID_BODY:
        umax = cfg.u0
    elif cfg.flow == FLOW_DEFORM:
        umax = cfg.u0 + cfg.k_def
    else:
        umax = cfg.u0
    return courant * h * cfg.R / umax


def as_ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))


# ---- reference structs (serbf_ref.h) ----
class NodesetStruct(C.Structure):
    _fields_ = [("Nnodes", C.c_int), ("n", C.c_int), ("x", c_double_p), ("y", c_double_p),
                ("z", c_double_p), ("lambda_", c_double_p), ("phi", c_double_p), ("idx", c_int_p),
                ("D", c_double_p)]


class LayersStruct(C.Structure):
    _fields_ = [("Nv", C.c_int), ("Ng", C.c_int), ("padded_Nv", C.c_int), ("Nvt", C.c_int),
                ("padded_Nvt", C.c_int), ("dh", C.c_double), ("htop", C.c_double), ("h", c_double_p),
                ("r", c_double_p), ("r_inv", c_double_p), ("p", c_double_p), ("rho", c_double_p)]


class HalosStruct(C.Structure):
    _fields_ = [("Nnbrs", C.c_int), ("halo_size_sum", C.c_int), ("nbr_halo_size_sum", C.c_int)] + [
        (k, c_int_p) for k in ("nbr_ranks", "halo_sizes", "halo_offsets", "nbr_halo_sizes",
                               "nbr_halo_offsets", "nbr_hid2pid", "nbr_hid2gid", "hid2pid", "hid2gid")]


class RbffdDMsStruct(C.Structure):
    _fields_ = [("part_Nnodes", C.c_int), ("n", C.c_int), ("gamma", C.c_double), ("idx", c_int_p),
                ("D", c_double_p), ("H", c_double_p), ("hDx", c_double_p), ("hDy", c_double_p),
                ("hDz", c_double_p), ("vDx", c_double_p), ("vDy", c_double_p), ("vDz", c_double_p),
                ("L", c_double_p), ("hDxp", c_double_p), ("hDyp", c_double_p), ("vDzp", C.c_double * 7)]


class PatchSVStruct(C.Structure):
    _fields_ = [("Ndim", C.c_int), ("data", c_double_p), ("halo_buff", c_double_p),
                ("nbr_halo_buff", c_double_p), ("request", C.c_void_p), ("status", C.c_void_p)]


class PartSVStruct(C.Structure):
    _fields_ = [("Ndim", C.c_int), ("data", c_double_p)]


class PatchStruct(C.Structure):
    _fields_ = [("Nnodes", C.c_int), ("part_Nnodes", C.c_int), ("nodeset", NodesetStruct),
                ("layers", LayersStruct), ("rbffd_DMs", RbffdDMsStruct), ("pid2gid", c_int_p),
                ("gid2pid", c_int_p), ("part_pids", c_int_p), ("halos", HalosStruct)] + [
        (k, C.POINTER(PatchSVStruct)) for k in ("SV_q_init", "SV_q_t", "SV_q_k")] + [
        (k, C.POINTER(PartSVStruct)) for k in ("SV_U", "SV_Up", "SV_Usph", "SV_F", "SV_F_k", "SV_3D_temp")]


class RtConfigStruct(C.Structure):
    _fields_ = [("nodeset_from_file", C.c_int), ("nodeset_file", C.c_char * 200), ("num_layers", C.c_int),
                ("model_height", C.c_double), ("stencil_size", C.c_int), ("TC", C.c_int),
                ("timestep_length", C.c_double), ("dt", C.c_double), ("nsteps", C.c_int),
                ("use_metis", C.c_int)]


class PhysConstantsStruct(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("R", "g", "p0", "cp", "Rd", "kappa", "T0", "rho0", "htop")]


class DomainsStruct(C.Structure):
    _fields_ = [("Nnodes", C.c_int), ("Nparts", C.c_int), ("global_nodeset", NodesetStruct),
                ("part_ids", c_int_p), ("part_sizes", c_int_p), ("part_start_ids", c_int_p)]


class TC2ParamsStruct(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("tau", "K", "u0", "w0", "h1", "h2", "htop")]


def lonlat(xyz):
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    lam = np.arctan2(y, x)
    lam = np.where(lam < 0, lam + 2 * np.pi, lam)
    phi = np.arctan2(z, np.sqrt(x * x + y * y))
    return lam, phi
