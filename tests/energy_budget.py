"""The energy budget of one COUPLED_ENERGY_EQS call (ppenergy.f:2-791),
evaluated in numpy from downloaded fields, for the oracle or the device.

What the reference's scheme conserves.  With theta = 1 the Newton system of
a cell is (ppenergy.f:286-345)
    Q-row:  FQ + dFQ dQ - dt rho kP c dE = 0
    E-row:  FR + dFR dQ + (1 + dt rho kP c) dE + dt ddivF = res,
where ddivF = sum_k EVAL_k dE(neighbour k) - EVAL_1_6 dE is the divergence of
the flux increment with the diffusion coefficients frozen (FILL_EVALUES,
ppenergy.f:984-1030), res the SOR residual of the eliminated row, and
dFQ + dFR = 1 - (dt / cv) heat dkA/dT (heat = STELLARINPUT / (rho kA): the
reference linearises the irradiation through kA while FQ holds it fixed).
Summing the rows, per cell
    dQ + dE + FQ + FR + dt ddivF - (dt / cv) heat dkA/dT dQ = res.      (1)
At the first Newton iteration Q = Q_old and E = ER_OLD, so
FQ + FR = dt (divF - STELLARINPUT).  Weighted with
    w = SURXB(i) DXA(i) DYA(j) SURZB(k) DZA(k),
which makes CALCULATEDIVF_WITHF's discrete divergence (ppenergy.f:825-878)
telescope exactly (the GITTER volumes VOLXB VOLYB VOLZB differ from w by up
to 1e-4, so they would leave a metric defect), the sums over the solved
region R become face sums over R's boundary:
    sum_R w (dQ + dE) = -dt Phi(F) - dt Phi(dF) + dt S + H + sum_R w res,   (2)
Phi(F) the net outflow of the flux through R's boundary faces, Phi(dF) that
of its frozen-coefficient increment (neighbours outside R have dE = 0:
identity rows), S = sum_R w STELLARINPUT, H = sum_R w (dt/cv) heat dkA/dT dQ.
The reference applies conv_parm = 0.1 of the correction (ppenergy.f:398-
408), so over one Newton iteration the energy in R changes by conv_parm
times (2).  R is i = 1 .. min(upperbnd - 1, nx - 1), k = 2 .. nz - 1 (BCXMIN
12): the cells above upperbnd and at i = nx, and the theta planes k = 1, nz,
are rows the solve leaves fixed and the energy boundary blocks then reset
(ppenergy.f:80-227); their change is reported apart.
"""
import numpy as np

# metric ids (include/ttgpu.h)
M_XXB, M_DXA, M_DYA, M_DZA = 1, 2, 6, 10
M_SURXA, M_SURXB, M_SURZA, M_SURZB = 22, 23, 34, 35
CONV_PARM = float(np.float32(0.1))  # REAL literal (ppenergy.f:49)


def _metric_fn(metric, info):
    """m -> function of the index array, from the declared lower bound."""
    def get(m):
        a = metric(m)
        lo = info(m)
        return lambda idx: a[np.asarray(idx) - lo]
    return get


def region(ub, nx, ny, nz):
    """Solved region R (BCXMIN 12) as a (nz, ny, nx) mask; ub is upperbnd
    over (k, j) interior columns."""
    i = np.arange(1, nx + 1)[None, None, :]
    k = np.arange(1, nz + 1)[:, None, None]
    top = np.minimum(ub[:, :, None] - 1, nx - 1)
    return (i <= top) & (k >= 2) & (k <= nz - 1)


def state_q_e(get):
    """Q = rho cv T and ER over the interior, wire order (k, j, i)."""
    inn = (slice(2, -2),) * 3
    return (get("RH") * get("CV") * get("T"))[inn], get("ER")[inn]


def budget(get, metric, info, nx, ny, nz, dt, q0, e0):
    """Terms of (1) and (2) after one call with a single Newton iteration.
    get(name) downloads a field (wire format), metric(m)/info(m) give a 1-D
    metric and its lower bound; q0, e0 are Q and ER before the call."""
    inn = (slice(2, -2),) * 3
    M = _metric_fn(metric, info)
    I = np.arange(1, nx + 1)[None, None, :]
    J = np.arange(1, ny + 1)[None, :, None]
    K = np.arange(1, nz + 1)[:, None, None]
    w = M(M_SURXB)(I) * M(M_DXA)(I) * M(M_DYA)(J) * M(M_SURZB)(K) * M(M_DZA)(K)
    ub = get("upperbnd")[2:nz + 2, 2:ny + 2]
    R = region(ub, nx, ny, nz)
    q1, e1 = state_q_e(get)
    rh, cv = get("RH")[inn], get("CV")[inn]
    c = CONV_PARM
    dQ = np.where(R, (q1 - q0) / c, 0.0)
    dE = np.where(R, get("delta_ER")[inn] / c, 0.0)  # delta_ER = ER - ER_OLD = conv_parm x the SOR solution in R

    def nbrs(x):
        p = np.zeros((nz + 2, ny + 2, nx + 2))
        p[1:-1, 1:-1, 1:-1] = x
        p[1:-1, 0, 1:-1] = x[:, -1]
        p[1:-1, -1, 1:-1] = x[:, 0]
        # i+1, i-1, j+1, j-1, k+1, k-1: the couplings of EVAL1..6
        return (p[1:-1, 1:-1, 2:], p[1:-1, 1:-1, :-2], p[1:-1, 2:, 1:-1], p[1:-1, :-2, 1:-1],
                p[2:, 1:-1, 1:-1], p[:-2, 1:-1, 1:-1])

    nb = nbrs(dE)
    ev = [get(f"EVAL{m}") for m in range(1, 7)]
    ddiv = sum(ev[m] * nb[m] for m in range(6)) - get("EVAL_1_6") * dE
    FQ, FR = get("FQ"), get("FR")
    sin = get("STELLARINPUT")[2:nz + 2, 2:ny + 2, :]
    heat = sin / (rh * get("xkapA")[inn])
    hterm = (dt / cv) * heat * get("dxkapA_dT")[inn] * dQ
    lhs = dQ + dE + FQ + FR + dt * ddiv - hterm
    # the E-row residual of the system solved (SOR_A..G, RHS_ER)
    A, B, C, D, E, F, G, RHS = (get(k) for k in ("SOR_A", "SOR_B", "SOR_C", "SOR_D", "SOR_E", "SOR_F", "SOR_G",
                                                 "RHS_ER"))
    ip, im, jp, jm, kp, km = nb
    res = A * dE + B * im + C * ip + D * jm + E * jp + F * km + G * kp - RHS
    scale = np.abs(dQ[R]).max() + np.abs(dE[R]).max()
    # the divergence of the flux after the solve (its last CALCULATEFLUX, at
    # ER_OLD for one Newton iteration) and its face sums over R's boundary
    fx, fy, fz = get("flux_x"), get("flux_y"), get("flux_z")  # (1:nx+1), (1:ny+1), (1:nz+1) on their axis
    sxa, sxb, dxa, xxb = M(M_SURXA), M(M_SURXB), M(M_DXA), M(M_XXB)
    sza, szb, dza, dya = M(M_SURZA), M(M_SURZB), M(M_DZA), M(M_DYA)
    divF = ((sxa(I + 1) / (sxb(I) * dxa(I))) * fx[:, :, 1:] - (sxa(I) / (sxb(I) * dxa(I))) * fx[:, :, :-1]) + \
           ((1.0 / (xxb(I) * szb(K) * dya(J))) * fy[:, 1:, :] - (1.0 / (xxb(I) * szb(K) * dya(J))) * fy[:, :-1, :]) + \
           ((sza(K + 1) / (xxb(I) * szb(K) * dza(K))) * fz[1:] - (sza(K) / (xxb(I) * szb(K) * dza(K))) * fz[:-1])
    # faces: area weights a with w divF = sum(+-a F); outflow = F (in_lo - in_hi)
    Rx = np.zeros((nz, ny, nx + 2), bool)
    Rx[:, :, 1:-1] = R                                   # cells i = 0..nx+1, faces i = 1..nx+1
    ax = (dya(J) * szb(K) * dza(K)) * sxa(np.arange(1, nx + 2)[None, None, :])
    phi_x = np.sum(ax * fx * (Rx[:, :, :-1].astype(float) - Rx[:, :, 1:]))
    Rz = np.zeros((nz + 2, ny, nx), bool)
    Rz[1:-1] = R
    az = (sxb(I) * dxa(I) * dya(J) / xxb(I)) * sza(np.arange(1, nz + 2)[:, None, None])
    phi_z = np.sum(az * fz * (Rz[:-1].astype(float) - Rz[1:]))
    Ry = np.concatenate([R[:, -1:, :], R, R[:, :1, :]], axis=1)   # periodic phi
    ay = (sxb(I) * dxa(I) * szb(K) * dza(K) / xxb(I))
    phi_y = np.sum(ay * fy * (Ry[:, :-1].astype(float) - Ry[:, 1:]))
    dphi = np.sum((w * ddiv)[R])                        # telescopes to the increment's outflow
    S = np.sum((w * sin)[R])
    H = np.sum((w * hterm)[R])
    dqe = np.sum((w * (dQ + dE))[R])
    return {
        "identity_err": float(np.abs(lhs - res)[R].max() / scale),   # (1), per cell, relative
        "sum_res": float(np.sum((w * res)[R])),
        "sum_abs_res_bound": float(np.max(w[R]) * np.sum(np.abs(res[R]))),
        "phi_faces": float(phi_x + phi_y + phi_z),
        "sum_w_divF": float(np.sum((w * divF)[R])),
        "sum_w_FQFR": float(np.sum((w * (FQ + FR))[R])),
        "phi_dF": dphi,
        "S": S, "H": H, "dt": dt,
        "sum_w_dQdE": dqe,
        "predicted": float(-dt * (phi_x + phi_y + phi_z) - dt * dphi + dt * S + H),
        "scale": float(np.sum(np.abs(w * (dQ + dE))[R]) + abs(dt * S)),
        "change_in_R": float(np.sum((w * ((q1 - q0) + (np.where(R, get("ER")[inn], 0) - np.where(R, e0, 0))))[R])),
        "change_outside_R": float(np.sum((w * ((q1 - q0) + (get("ER")[inn] - e0)))[~R])),
    }
