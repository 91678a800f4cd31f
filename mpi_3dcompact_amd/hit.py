"""Decaying compressible HIT workload: options deck, mesh metrics and initial field.

The initial velocity follows the Rogallo construction of the reference's HIT
generator (tests/hit/generate_init_hit.f90:246-363): random phases phi1, phi2,
phi3 per wave vector, a = sqrt(E/(2 pi k^2)) e^{i phi1} cos phi3,
b = sqrt(E/(2 pi k^2)) e^{i phi2} sin phi3, the solenoidal projection onto the
plane normal to k, odd-ball (Nyquist) modes and the mean removed, inverse FFT.
The spectrum is E(k) = (k/k0)^4 exp(-2 (k/k0)^2).  Phases come from numpy's
PCG64 generator with an explicit seed (the Fortran LCG shuffle is not
reproduced).  The field is scaled so that the turbulent Mach number
M_t = sqrt(<u.u>)/<c> equals the requested value (the definition of
tests/hit/Samtaney_D4/MPI_3DCompact.cpp:455,485-486,535-537), with rho = 1 and
p = c0^2/gamma: c0 = 1 by default (T = 1), c0^2 = 11.11111111 in the
reference's Samtaney D4 case (tests/hit/Samtaney_D4/MPI_3DCompact.cpp:230-231).

The mesh is the HIT identity map x = 2 pi xi (tests/hit/Samtaney_D4/
AlgebraicSingleBlockMesh.hpp:151-156), whose conservative metrics are
J = (2 pi)^-3, J11 = J22 = J33 = 1/(2 pi) and zero off-diagonal terms.
"""
import numpy as np

GAMMA = 1.4


def hit_velocity(N, k0=4.0, Mt=0.3, seed=0, c0=1.0):
    """Solenoidal random velocity (U, V, W), each of shape (N, N, N) indexed
    [k][j][i], scaled to M_t at the uniform sound speed c0."""
    rng = np.random.default_rng(seed)
    kf = np.fft.fftfreq(N, d=1.0 / N)          # integer wavenumbers, FFT order
    kr = np.fft.rfftfreq(N, d=1.0 / N)
    kz, ky, kx = np.meshgrid(kf, kf, kr, indexing="ij")
    k = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    kxy = np.sqrt(kx ** 2 + ky ** 2)
    shape = k.shape
    ph1 = 2 * np.pi * (rng.random(shape) - 0.5)
    ph2 = 2 * np.pi * (rng.random(shape) - 0.5)
    ph3 = 2 * np.pi * (rng.random(shape) - 0.5)
    kmin = 1e-10
    E = (k / k0) ** 4 * np.exp(-2.0 * (k / k0) ** 2)
    amp = np.where(k > kmin, np.sqrt(E / (2 * np.pi * np.maximum(k, kmin) ** 2)), 0.0)
    a = amp * np.exp(1j * ph1) * np.cos(ph3)
    b = amp * np.exp(1j * ph2) * np.sin(ph3)
    safe_k = np.maximum(k, kmin)
    safe_kxy = np.maximum(kxy, kmin)
    on_axis = kxy <= kmin
    ux = np.where(on_axis, (a + b) / np.sqrt(2.0), (a * k * ky + b * kx * kz) / (safe_k * safe_kxy))
    uy = np.where(on_axis, (b - a) / np.sqrt(2.0), (b * ky * kz - a * k * kx) / (safe_k * safe_kxy))
    uz = -b * kxy / safe_k
    out = []
    for uk in (ux, uy, uz):
        uk = np.where(k > kmin, uk, 0.0)
        # remove odd-ball wave numbers (the Nyquist planes) and the mean
        uk[N // 2, :, :] = 0.0
        uk[:, N // 2, :] = 0.0
        uk[:, :, -1] = 0.0
        uk[0, 0, 0] = 0.0
        out.append(np.fft.irfftn(uk, s=(N, N, N), axes=(0, 1, 2)))
    U, V, W = out
    urms = np.sqrt(np.mean(U ** 2 + V ** 2 + W ** 2))
    scale = Mt * c0 / urms  # the mean speed of sound is c0
    return U * scale, V * scale, W * scale


def hit_initial_condition(N, k0=4.0, Mt=0.3, seed=0, c0=1.0):
    U, V, W = hit_velocity(N, k0, Mt, seed, c0)
    rho = np.ones((N, N, N))
    p = np.full((N, N, N), c0 * c0 / GAMMA)
    return dict(rho=rho, U=U, V=V, W=W, p=p)


def identity_metrics(N):
    """J, J11..J33 of x = 2 pi xi (10 arrays, shape (N, N, N))."""
    L = 2 * np.pi
    one = np.ones((N, N, N))
    zero = np.zeros((N, N, N))
    J = one / L ** 3
    d = one / L
    return [J, d, zero, zero, zero, d, zero, zero, zero, d]


def options_text(N, fd="PADE6", filt="COMPACT10", alpha_f=0.49, rk="TVDRK3", cfl=0.5,
                 mu_ref=1.0 / 1600.0, les="NOMODEL", les_avg="NONE", max_steps=100000,
                 filter_step=1, substep_filtering=False, dsm_filter=1, use_taukk=True,
                 coeff_range=False, ts_type="CONST_CFL", dt=None, restart_file=None,
                 dump_step=1000000, disp=None, closure_file=None, closure_fallback=True,
                 faces=None, sponge=None, closure_fallback_fraction=None, stats=None, dims=None,
                 sponge_avg_file=None, closure_features=None):
    """MPI_3DCompact.in text for a periodic HIT box of N^3 points (restart_file:
    start from that SolutionDump, which also sets the grid; disp: 3x3
    PERIODICDISP rows, default the 2 pi box; stats: dict velocity, thermo (bools),
    avg (STATSAVGTYPE name), interval; dims: (Nx, Ny, Nz) instead of N^3;
    sponge_avg_file: SpongeAvgDump file the sponge averages start from)."""
    Nx, Ny, Nz = dims if dims is not None else (N, N, N)
    fds = fd if isinstance(fd, (list, tuple)) else (fd, fd, fd)
    lines = ["[RESTART]"]
    if restart_file:
        lines += ["FROMRESTART = TRUE", "ONLYGRIDFROMRESTART = FALSE", f"FILENAME = {restart_file}"]
    else:
        lines += ["FROMRESTART = FALSE", "ONLYGRIDFROMRESTART = FALSE"]
    fk = face_kinds(faces)
    lines += [
        "SPONGEAVGFROMRESTART = " + ("TRUE" if sponge_avg_file else "FALSE"),
    ] + ([f"SPONGEAVGFILENAME = {sponge_avg_file}"] if sponge_avg_file else []) + [
        "[TIMESTEPPING]", f"TSTYPE = {ts_type}",
        (f"CFL = {cfl!r}" if ts_type == "CONST_CFL" else f"DT = {dt!r}"),
        "MAXTIME = 1000000.0", f"MAXTIMESTEP = {max_steps}", f"FILTERSTEP = {filter_step}",
        f"SUBSTEPFILTERING = {'TRUE' if substep_filtering else 'FALSE'}",
        "CHECKSTEP = 1000000", f"DUMPSTEP = {dump_step}",
        "[STATS]", f"STATSAVGTYPE = {(stats or {}).get('avg', 'NONE')}",
        f"VELOCITYSTATS = {'TRUE' if (stats or {}).get('velocity') else 'FALSE'}",
        f"THERMOSTATS = {'TRUE' if (stats or {}).get('thermo') else 'FALSE'}",
        "VELOCITYSTATSFROMRESTART = " + ("TRUE" if (stats or {}).get("velocity_file") else "FALSE"),
        "THERMOSTATSFROMRESTART = " + ("TRUE" if (stats or {}).get("thermo_file") else "FALSE"),
    ] + ([f"VELOCITYSTATSFILENAME = {stats['velocity_file']}"] if (stats or {}).get("velocity_file") else []) + (
        [f"THERMOSTATSFILENAME = {stats['thermo_file']}"] if (stats or {}).get("thermo_file") else []) + [
        f"INTERVAL = {int((stats or {}).get('interval', 5))}",
        "[BC]"]
    for d, a in enumerate("XYZ"):
        per = fk[2 * d] in _PERIODIC
        lines.append(f"BC{a}TYPE = {'PERIODIC_SOLVE' if per else 'DIRICHLET_SOLVE'}")
        lines += [f"BC{a}0 = {fk[2 * d]}", f"BC{a}1 = {fk[2 * d + 1]}"]
    if sponge is not None:
        lines.append("[SPONGE]")
        lines += [f"{k} = {v}" for k, v in sponge.items()]
        lines.append("[BC]")
    if disp is None:
        disp = 2 * np.pi * np.eye(3)
    for i, a in enumerate("XYZ"):
        if fk[2 * i] != "PERIODIC":
            continue
        for j, b in enumerate("XYZ"):
            lines.append(f"PERIODICDISP{a}_{b} = {float(disp[i][j])!r}")
    lines += [
        "[DOMAIN]"] + ([] if restart_file else [f"NX = {Nx}", f"NY = {Ny}", f"NZ = {Nz}"]) + [
        "[PENCILDECOMP]", "PROW = 0", "PCOL = 0",
        "[SOLVER]", f"ALPHAF = {alpha_f!r}", f"MU_REF = {mu_ref!r}", "USETIMING = FALSE",
        f"XFDTYPE = {fds[0]}", f"YFDTYPE = {fds[1]}", f"ZFDTYPE = {fds[2]}",
        f"FILTERTYPE = {filt}", f"RKTYPE = {rk}",
        "[LES]", f"LESMODEL = {les}", f"LESAVERAGING = {les_avg}", f"DSMTESTFILTER = {dsm_filter}",
        f"DSMUSETAUKK = {'TRUE' if use_taukk else 'FALSE'}",
        f"DSMCOEFFRANGE = {'TRUE' if coeff_range else 'FALSE'}",
        f"CLOSUREFALLBACK = {'TRUE' if closure_fallback else 'FALSE'}",
    ] + ([f"CLOSUREFILE = {closure_file}"] if closure_file else []) + (
        [f"CLOSUREFALLBACKFRACTION = {closure_fallback_fraction!r}"] if closure_fallback_fraction is not None else []) + (
        [f"CLOSUREFEATURES = {closure_features}"] if closure_features else [])
    return "\n".join(lines) + "\n"


def ensemble_members(count=64, seed=2024, sizes=(32, 48, 64, 96, 128), mt=(0.1, 0.4),
                     mu=(1.0 / 3200.0, 1.0 / 800.0)):
    """Members of a decaying-HIT ensemble: N drawn uniformly from sizes, M_t
    uniform in mt, MU_REF uniform in mu, and an initial-field seed each, from
    one seeded generator.  Returns a list of dicts N, Mt, mu, seed."""
    r = np.random.default_rng(seed)
    out = []
    for m in range(count):
        out.append(dict(N=int(r.choice(sizes)), Mt=float(r.uniform(*mt)), mu=float(r.uniform(*mu)),
                        seed=int(r.integers(0, 2 ** 31 - 1))))
    return out


FACE_NAMES = ("X0", "X1", "Y0", "Y1", "Z0", "Z1")
_PERIODIC = ("PERIODIC", "INTERNALLY_PERIODIC")  # faces of a periodic (PERIODIC_SOLVE) direction
_FACE = {"PERIODIC": 0, "INTERNALLY_PERIODIC": 1, "SPONGE": 2, "ADIABATIC_WALL": 3, "CONST_T_WALL": 4,
         "MOVING_ADIABATIC_WALL": 5, "INLET": 6}


def face_kinds(faces):
    """faces: {"Y0": "ADIABATIC_WALL", ...} (missing faces are PERIODIC) -> six kind names."""
    faces = faces or {}
    out = [faces.get(f, "PERIODIC") for f in FACE_NAMES]
    for d in range(3):
        if (out[2 * d] in _PERIODIC) != (out[2 * d + 1] in _PERIODIC):
            raise ValueError("both faces of a direction must be periodic or neither")
    return out


def box_mesh(N, faces=None, L=2 * np.pi):
    """Coordinates x = L xi of the HIT box, xi in [0, 1) periodic or [0, 1] otherwise ((N, N, N) arrays)."""
    fk = face_kinds(faces)
    axes = [np.arange(N) / (N if fk[2 * d] in _PERIODIC else N - 1) * L for d in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return x, y, z


_FD = {"CD2": 0, "PADE6": 1, "PENTA10": 2}
_FILT = {"COMPACT8": 0, "COMPACT10": 1}
_RK = {"TVDRK3": 0, "RK4": 1, "KENRK4": 2, "LSLDDRK4": 3}
_LES = {"NOMODEL": 0, "NONE": 0, "VREMAN": 1, "DSM": 2}
_AVG = {"NONE": 0, "XI1_AVG": 1, "XI2_AVG": 2, "XI3_AVG": 3, "LOCAL": 4, "GLOBAL": 5}


def oracle_config(N, fd="PADE6", filt="COMPACT10", alpha_f=0.49, rk="TVDRK3", cfl=0.5,
                  mu_ref=1.0 / 1600.0, les="NOMODEL", les_avg="NONE", filter_step=1,
                  substep_filtering=False, dsm_filter=1, use_taukk=True, coeff_range=False,
                  ts_type="CONST_CFL", dt=None, faces=None, stats=None, dims=None, **_):
    """The same case as options_text(), as the field dict of the CPU oracle's config."""
    fds = fd if isinstance(fd, (list, tuple)) else (fd, fd, fd)
    fk = face_kinds(faces)
    bc = tuple(0 if fk[2 * d] in _PERIODIC else 1 for d in range(3))
    n3 = dims if dims is not None else (N, N, N)
    sp = [1.0 / n3[d] if bc[d] == 0 else 1.0 / (n3[d] - 1) for d in range(3)]
    return dict(Nx=n3[0], Ny=n3[1], Nz=n3[2], bc=bc, face=tuple(_FACE[k] for k in fk), fd=tuple(_FD[f] for f in fds),
                filter=_FILT[filt], alphaF=alpha_f, rk=_RK[rk], les=_LES[les],
                les_avg=_AVG[les_avg], dsm_test_filter=dsm_filter, dsm_use_taukk=int(use_taukk),
                dsm_coeff_range=int(coeff_range), mu_ref=mu_ref,
                ts_type=1 if ts_type == "CONST_CFL" else 0, cfl=cfl if cfl else 0.0,
                dt=dt if dt else 0.0, filter_step=filter_step,
                substep_filtering=int(substep_filtering), dx=sp[0], dy=sp[1], dz=sp[2],
                timestep0=0, time0=0.0,
                velocity_stats=int(bool((stats or {}).get("velocity"))),
                thermo_stats=int(bool((stats or {}).get("thermo"))),
                stats_avg=_AVG[(stats or {}).get("avg", "NONE")],
                stats_interval=int((stats or {}).get("interval", 5)))
