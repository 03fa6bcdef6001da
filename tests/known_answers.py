"""One ionic step of each model restated a second time, in Python, for the
oracle's known-answer tests (test infrastructure).

These restatements are written from the reference's kernels
(cpuwave2D/model/*.py), independently of oracle/fw_oracle.c, and take their
parameters from tests/golden/reference_tables.json (read from the
reference's source with ast), not from the oracle's tables.  A constant, sign
or index typo in either restatement makes the two disagree.

* Aliev-Panfilov, Barkley, Mitchell-Schaeffer: exact rational arithmetic
  (fractions.Fraction) on the binary values of the inputs, so the oracle's
  FP64 result must lie within a few rounding errors of the exact one.
* Fenton-Karma, Bueno-Orovio: exact arithmetic except tanh, whose argument
  is formed in FP64 as the reference forms it and evaluated by math.tanh.
* Luo-Rudy 91, TP06, Courtemanche: scalar FP64 in the reference's operation
  order with the host libm (math.exp/log/sqrt), `x**n` as Python's pow.

Each function returns (du, new state) where du is the kernel's change of
u_new (u_new += du) and the state is a dict in the kernel's slot names.
"""
from __future__ import annotations

import math
from fractions import Fraction as Q


def _f(x) -> Q:
    return Q(x)


# ---------------------------------------------------------------- exact models
def aliev_panfilov(u, v, dt, p):
    # aliev_panfilov_2d.py:133-170 (calc_v), :199-202
    u, v, dt = _f(u), _f(v), _f(dt)
    a, k, eap, mu_1, mu_2 = (_f(p[n]) for n in ("a", "k", "eap", "mu_1", "mu_2"))
    v = v + (-dt * (eap + (mu_1 * v) / (mu_2 + u)) * (v + k * u * (u - a - 1)))
    du = dt * (-k * u * (u - a) * (u - 1) - u * v)
    return du, {"v": v}


def barkley(u, v, dt, p):
    # barkley_2d.py:113-136 (calc_v), :164-166
    u, v, dt = _f(u), _f(v), _f(dt)
    a, b, eap = (_f(p[n]) for n in ("a", "b", "eap"))
    v = v + dt * (u - v)
    du = dt * (u * (1 - u) * (u - (v + b) / a)) / eap
    return du, {"v": v}


def mitchell_schaeffer(u, h, dt, p):
    # mitchell_schaeffer_2d.py:102-217: calc_h, calc_J_in, calc_J_out
    u, h, dt = _f(u), _f(h), _f(dt)
    tau_close, tau_open, tau_in, tau_out, u_gate = (
        _f(p[n]) for n in ("tau_close", "tau_open", "tau_in", "tau_out", "u_gate"))
    h = h + (dt * (1 - h) / tau_open if u < u_gate else dt * (-h) / tau_close)
    j_in = h * (u ** 2 * (1 - u)) / tau_in
    j_out = -u / tau_out
    return dt * (j_in + j_out), {"h": h}


def _tanh_q(x_float: float) -> Q:
    return Q(math.tanh(x_float))


def fenton_karma(u, v, w, dt, p):
    # fenton_karma_2d.py:143-328; J_si takes the updated v (:326)
    uf = u
    u, v, w, dt = _f(u), _f(v), _f(w), _f(dt)
    g = {n: _f(p[n]) for n in ("tau_d", "tau_o", "tau_r", "tau_si", "tau_v_m", "tau_v_p",
                                "tau_w_m", "tau_w_p", "k", "u_c", "uc_si")}
    H1 = 1 if (g["u_c"] - u) >= 0 else 0
    H2 = 1 if (u - g["u_c"]) >= 0 else 0
    v = v + dt * (H1 * (1 - v) / g["tau_v_m"] - H2 * v / g["tau_v_p"])
    w = w + dt * (H1 * (1 - w) / g["tau_w_m"] - H2 * w / g["tau_w_p"])
    H = 1 if (u - g["u_c"]) >= 0 else 0
    j_fi = -(v * H * (1 - u) * (u - g["u_c"])) / g["tau_d"]
    j_so = u * H1 / g["tau_o"] + H2 / g["tau_r"]
    th = _tanh_q(p["k"] * (uf - p["uc_si"]))
    j_si = -v * (1 + th) / (2 * g["tau_si"])
    return dt * (-j_fi - j_so - j_si), {"v": v, "w": w}


def bueno_orovio(u, v, w, s, dt, p):
    # bueno_orovio_2d.py:183-476
    uf = u
    u, v, w, s, dt = _f(u), _f(v), _f(w), _f(s), _f(dt)
    g = {n: _f(x) for n, x in p.items()}
    v_inf = 1 if u < g["theta_v_m"] else 0
    tau_v_m = g["tau_v1_m"] if (u - g["theta_v_m"]) < 0 else g["tau_v2_m"]
    v = v + dt * ((v_inf - v) / tau_v_m if (u - g["theta_v"]) < 0 else -v / g["tau_v_p"])
    w_inf = 1 - u / g["tau_w_inf"] if (u - g["theta_o"]) < 0 else g["w_inf_"]
    tau_w_m = g["tau_w1_m"] + (g["tau_w2_m"] - g["tau_w1_m"]) * \
        (1 + _tanh_q(p["k_w_m"] * (uf - p["u_w_m"]))) / 2
    w = w + dt * ((w_inf - w) / tau_w_m if (u - g["theta_w"]) < 0 else -w / g["tau_w_p"])
    tau_s = g["tau_s1"] if (u - g["theta_w"]) < 0 else g["tau_s2"]
    s = s + dt * ((1 + _tanh_q(p["k_s"] * (uf - p["u_s"]))) / 2 - s) / tau_s
    H = 1 if (u - g["theta_v"]) >= 0 else 0
    j_fi = -v * H * (u - g["theta_v"]) * (g["u_u"] - u) / g["tau_fi"]
    Hw = 1 if (u - g["theta_w"]) >= 0 else 0
    tau_o = g["tau_o1"] if (u - g["theta_o"]) < 0 else g["tau_o2"]
    tau_so = g["tau_so1"] + (g["tau_so2"] - g["tau_so1"]) * \
        (1 + _tanh_q(p["k_so"] * (uf - p["u_so"]))) / 2
    j_so = (u - g["u_o"]) * (1 - Hw) / tau_o + Hw / tau_so
    j_si = -Hw * w * s / g["tau_si"]
    return dt * (-j_fi - j_so - j_si), {"v": v, "w": w, "s": s}


# ---------------------------------------------------------------- TP06
def tp06(u, st, dt, p):
    """tp06_2d.py:242-1103: the helpers calc_* and ionic_kernel_2d.  st: the
    kernel's slots cai, casr, cass, nai, Ki, m, h, j_, xr1, xr2, xs, r, s, d,
    f, f2, fcass, rr, oo."""
    exp, log, sqrt = math.exp, math.log, math.sqrt
    P = p
    ko, cao, nao = P["ko"], P["cao"], P["nao"]
    Vc, Vsr, Vss = P["Vc"], P["Vsr"], P["Vss"]
    R, F, T, RTONF = P["R"], P["F"], P["T"], P["RTONF"]
    CAP = P["CAPACITANCE"]
    inverseVcF2 = 1. / (2 * Vc * F)
    inverseVcF = 1. / (Vc * F)
    inversevssF2 = 1. / (2 * Vss * F)
    cai, casr, cass, nai, Ki = st["cai"], st["casr"], st["cass"], st["nai"], st["Ki"]
    Ek = RTONF * (log((ko / Ki)))
    Ena = RTONF * (log((nao / nai)))
    Eks = RTONF * (log((ko + P["pKNa"] * nao) / (Ki + P["pKNa"] * nai)))
    Eca = 0.5 * RTONF * (log((cao / cai)))
    out = dict(st)
    # calc_ina
    m, h, j = st["m"], st["h"], st["j_"]
    alpha_m = 1. / (1. + exp((-60. - u) / 5.))
    beta_m = 0.1 / (1. + exp((u + 35.) / 5.)) + 0.10 / (1. + exp((u - 50.) / 200.))
    tau_m = alpha_m * beta_m
    m_inf = 1. / ((1. + exp((-56.86 - u) / 9.03)) * (1. + exp((-56.86 - u) / 9.03)))
    if u >= -40.:
        alpha_h = 0.
        beta_h = 0.77 / (0.13 * (1. + exp(-(u + 10.66) / 11.1)))
    else:
        alpha_h = 0.057 * exp(-(u + 80.) / 6.8)
        beta_h = 2.7 * exp(0.079 * u) + (3.1e5) * exp(0.3485 * u)
    tau_h = 1.0 / (alpha_h + beta_h)
    h_inf = 1. / ((1. + exp((u + 71.55) / 7.43)) * (1. + exp((u + 71.55) / 7.43)))
    if u >= -40.:
        alpha_j = 0.
        beta_j = 0.6 * exp((0.057) * u) / (1. + exp(-0.1 * (u + 32.)))
    else:
        alpha_j = ((-2.5428e4) * exp(0.2444 * u) - (6.948e-6) * exp(-0.04391 * u)) * \
            (u + 37.78) / (1. + exp(0.311 * (u + 79.23)))
        beta_j = 0.02424 * exp(-0.01052 * u) / (1. + exp(-0.1378 * (u + 40.14)))
    tau_j = 1.0 / (alpha_j + beta_j)
    j_inf = h_inf
    m = m_inf - (m_inf - m) * exp(-dt / tau_m)
    h = h_inf - (h_inf - h) * exp(-dt / tau_h)
    j = j_inf - (j_inf - j) * exp(-dt / tau_j)
    ina = P["gna"] * m * m * m * h * j * (u - Ena)
    out.update(m=m, h=h, j_=j)
    # calc_ical
    d, f, f2, fcass = st["d"], st["f"], st["f2"], st["fcass"]
    d_inf = 1. / (1. + exp((-8 - u) / 7.5))
    Ad = 1.4 / (1. + exp((-35 - u) / 13)) + 0.25
    Bd = 1.4 / (1. + exp((u + 5) / 5))
    Cd = 1. / (1. + exp((50 - u) / 20))
    tau_d = Ad * Bd + Cd
    f_inf = 1. / (1. + exp((u + 20) / 7))
    Af = 1102.5 * exp(-(u + 27) * (u + 27) / 225)
    Bf = 200. / (1 + exp((13 - u) / 10.))
    Cf = (180. / (1 + exp((u + 30) / 10))) + 20
    tau_f = Af + Bf + Cf
    f2_inf = 0.67 / (1. + exp((u + 35) / 7)) + 0.33
    Af2 = 600 * exp(-(u + 25) * (u + 25) / 170)
    Bf2 = 31 / (1. + exp((25 - u) / 10))
    Cf2 = 16 / (1. + exp((u + 30) / 10))
    tau_f2 = Af2 + Bf2 + Cf2
    fcass_inf = 0.6 / (1 + (cass / 0.05) * (cass / 0.05)) + 0.4
    tau_fcass = 80. / (1 + (cass / 0.05) * (cass / 0.05)) + 2.
    d = d_inf - (d_inf - d) * exp(-dt / tau_d)
    f = f_inf - (f_inf - f) * exp(-dt / tau_f)
    f2 = f2_inf - (f2_inf - f2) * exp(-dt / tau_f2)
    fcass = fcass_inf - (fcass_inf - fcass) * exp(-dt / tau_fcass)
    ical = P["gcal"] * d * f * f2 * fcass * 4 * (u - 15) * (F * F / (R * T)) * \
        (0.25 * exp(2 * (u - 15) * F / (R * T)) * cass - cao) / \
        (exp(2 * (u - 15) * F / (R * T)) - 1.)
    out.update(d=d, f=f, f2=f2, fcass=fcass)
    # calc_ito
    r, s = st["r"], st["s"]
    r_inf = 1. / (1. + exp((20 - u) / 6.))
    s_inf = 1. / (1. + exp((u + 20) / 5.))
    tau_r = 9.5 * exp(-(u + 40.) * (u + 40.) / 1800.) + 0.8
    tau_s = 85. * exp(-(u + 45.) * (u + 45.) / 320.) + 5. / (1. + exp((u - 20.) / 5.)) + 3.
    s = s_inf - (s_inf - s) * exp(-dt / tau_s)
    r = r_inf - (r_inf - r) * exp(-dt / tau_r)
    ito = P["gto"] * r * s * (u - Ek)
    out.update(r=r, s=s)
    # calc_ikr
    xr1, xr2 = st["xr1"], st["xr2"]
    xr1_inf = 1. / (1. + exp((-26. - u) / 7.))
    axr1 = 450. / (1. + exp((-45. - u) / 10.))
    bxr1 = 6. / (1. + exp((u - (-30.)) / 11.5))
    tau_xr1 = axr1 * bxr1
    xr2_inf = 1. / (1. + exp((u - (-88.)) / 24.))
    axr2 = 3. / (1. + exp((-60. - u) / 20.))
    bxr2 = 1.12 / (1. + exp((u - 60.) / 20.))
    tau_xr2 = axr2 * bxr2
    xr1 = xr1_inf - (xr1_inf - xr1) * exp(-dt / tau_xr1)
    xr2 = xr2_inf - (xr2_inf - xr2) * exp(-dt / tau_xr2)
    ikr = P["gkr"] * sqrt(ko / 5.4) * xr1 * xr2 * (u - Ek)
    out.update(xr1=xr1, xr2=xr2)
    # calc_iks
    xs = st["xs"]
    Axs = (1400. / (sqrt(1. + exp((5. - u) / 6))))
    Bxs = (1. / (1. + exp((u - 35.) / 15.)))
    tau_xs = Axs * Bxs + 80
    xs_inf = 1. / (1. + exp((-5. - u) / 14.))
    xs = xs_inf - (xs_inf - xs) * exp(-dt / tau_xs)
    iks = P["gks"] * xs * xs * (u - Eks)
    out.update(xs=xs)
    # calc_ik1
    ak1 = 0.1 / (1. + exp(0.06 * (u - Ek - 200)))
    bk1 = (3. * exp(0.0002 * (u - Ek + 100)) + exp(0.1 * (u - Ek - 10))) / \
        (1. + exp(-0.5 * (u - Ek)))
    rec_iK1 = ak1 / (ak1 + bk1)
    ik1 = P["gk1"] * rec_iK1 * (u - Ek)
    # calc_inaca
    KmNai, KmCa, n_ = P["KmNai"], P["KmCa"], P["n_"]
    inaca = P["knaca"] * (1. / (KmNai * KmNai * KmNai + nao * nao * nao)) * \
        (1. / (KmCa + cao)) * (1. / (1 + P["ksat"] * exp((n_ - 1) * u * F / (R * T)))) * \
        (exp(n_ * u * F / (R * T)) * nai * nai * nai * cao -
         exp((n_ - 1) * u * F / (R * T)) * nao * nao * nao * cai * 2.5)
    # calc_inak
    rec_iNaK = (1. / (1. + 0.1245 * exp(-0.1 * u * F / (R * T)) + 0.0353 * exp(-u * F / (R * T))))
    inak = P["knak"] * (ko / (ko + P["KmK"])) * (nai / (nai + P["KmNa"])) * rec_iNaK
    ipca = P["gpca"] * cai / (P["KpCa"] + cai)
    rec_ipK = 1. / (1. + exp((25 - u) / 5.98))
    ipk = P["gpk"] * rec_ipK * (u - Ek)
    ibna = P["gbna"] * (u - Ena)
    ibca = P["gbca"] * (u - Eca)
    # calc_irel
    rr = st["rr"]
    maxsr, minsr, EC = P["maxsr"], P["minsr"], P["EC"]
    kCaSR = maxsr - ((maxsr - minsr) / (1 + (EC / casr) * (EC / casr)))
    k1_ = P["k1_"] / kCaSR
    k2_ = P["k2_"] * kCaSR
    drr = P["k4"] * (1 - rr) - k2_ * cass * rr
    rr = rr + dt * drr
    oo = k1_ * cass * cass * rr / (P["k3"] + k1_ * cass * cass)
    irel = P["Vrel"] * oo * (casr - cass)
    out.update(rr=rr, oo=oo)
    ileak = P["Vleak"] * (casr - cai)
    iup = P["Vmaxup"] / (1. + ((P["Kup"] * P["Kup"]) / (cai * cai)))
    ixfer = P["Vxfer"] * (cass - cai)
    # calc_casr
    bufsr, Kbufsr = P["Bufsr"], P["Kbufsr"]
    CaCSQN = bufsr * casr / (casr + Kbufsr)
    dCaSR = dt * (iup - irel - ileak)
    bjsr = bufsr - CaCSQN - dCaSR - casr + Kbufsr
    cjsr = Kbufsr * (CaCSQN + dCaSR + casr)
    out["casr"] = (sqrt(bjsr * bjsr + 4 * cjsr) - bjsr) / 2
    # calc_cass
    bufss, Kbufss = P["Bufss"], P["Kbufss"]
    CaSSBuf = bufss * cass / (cass + Kbufss)
    dCaSS = dt * (-ixfer * (Vc / Vss) + irel * (Vsr / Vss) + (-ical * inversevssF2 * CAP))
    bcss = bufss - CaSSBuf - dCaSS - cass + Kbufss
    ccss = Kbufss * (CaSSBuf + dCaSS + cass)
    out["cass"] = (sqrt(bcss * bcss + 4 * ccss) - bcss) / 2
    # calc_cai returns (new, cai); `cai, cai = ...` keeps the old value
    out["cai"] = cai
    out["nai"] = nai + dt * (-(ina + ibna + 3 * inak + 3 * inaca) * inverseVcF * CAP)
    out["Ki"] = Ki + dt * (-(ik1 + ito + ikr + iks - 2 * inak + ipk) * inverseVcF * CAP)
    du = -(dt * (ikr + iks + ik1 + ito + ina + ibna + ical + ibca + inak + inaca + ipca + ipk))
    return du, out


def tp06_cai_update(cai, dt, p, ibca, ipca, inaca, iup, ileak, ixfer):
    """calc_cai's first return value (tp06_2d.py:875-920), which the kernel
    discards: the closed-form buffered cytosolic calcium."""
    bufc, Kbufc = p["Bufc"], p["Kbufc"]
    inverseVcF2 = 1. / (2 * p["Vc"] * p["F"])
    CaCBuf = bufc * cai / (cai + Kbufc)
    dCai = dt * ((-(ibca + ipca - 2 * inaca) * inverseVcF2 * p["CAPACITANCE"]) -
                 (iup - ileak) * (p["Vsr"] / p["Vc"]) + ixfer)
    bc = bufc - CaCBuf - dCai - cai + Kbufc
    cc = Kbufc * (CaCBuf + dCai + cai)
    return (math.sqrt(bc * bc + 4 * cc) - bc) / 2


# ---------------------------------------------------------------- Luo-Rudy 91
def _lr_gate(var, dt, alpha, beta):
    # luo_rudy91_2d.py calc_gating_var: forward Euler toward alpha/(alpha+beta)
    tau = 1.0 / (alpha + beta)
    inf = alpha / (alpha + beta)
    return var + dt * (inf - var) / tau


def luo_rudy91(u, st, dt, p):
    """luo_rudy91_2d.py:158-510.  st: the kernel's slots m, h, j_, d, f, x,
    cai.  E_Na and E_K1 take the scalar parameters nai and ki (:478, :496)."""
    exp, log, sqrt = math.exp, math.log, math.sqrt
    R, T, F = p["R"], p["T"], p["F"]
    ko, ki, nai, nao = p["ko"], p["ki"], p["nai"], p["nao"]
    out = dict(st)
    E_Na = R * T / F * log(nao / nai)
    # calc_ina
    alpha_h = beta_h = beta_J = alpha_J = 0
    if u >= -40.0:
        beta_h = 1.0 / (0.13 * (1 + exp((u + 10.66) / -11.1)))
        beta_J = 0.3 * exp(-2.535 * 1e-07 * u) / (1 + exp(-0.1 * (u + 32)))
    else:
        alpha_h = 0.135 * exp((80 + u) / -6.8)
        beta_h = 3.56 * exp(0.079 * u) + 3.1 * 100000.0 * exp(0.35 * u)
        beta_J = 0.1212 * exp(-0.01052 * u) / (1 + exp(-0.1378 * (u + 40.14)))
        alpha_J = (-1.2714 * 100000.0 * exp(0.2444 * u) - 3.474 * 1e-05 * exp(-0.04391 * u)) * \
            (u + 37.78) / (1 + exp(0.311 * (u + 79.23)))
    alpha_m = 0.32 * (u + 47.13) / (1 - exp(-0.1 * (u + 47.13)))
    beta_m = 0.08 * exp(-u / 11)
    m = _lr_gate(st["m"], dt, alpha_m, beta_m)
    h = _lr_gate(st["h"], dt, alpha_h, beta_h)
    j = _lr_gate(st["j_"], dt, alpha_J, beta_J)
    ina = p["gna"] * m * m * m * h * j * (u - E_Na)
    out.update(m=m, h=h, j_=j)
    # calc_isk
    cai = st["cai"]
    E_Si = 7.7 - 13.0287 * log(cai)
    I_Si = p["gsi"] * st["d"] * st["f"] * (u - E_Si)
    alpha_d = 0.095 * exp(-0.01 * (u - 5)) / (1 + exp(-0.072 * (u - 5)))
    beta_d = 0.07 * exp(-0.017 * (u + 44)) / (1 + exp(0.05 * (u + 44)))
    alpha_f = 0.012 * exp(-0.008 * (u + 28)) / (1 + exp(0.15 * (u + 28)))
    beta_f = 0.0065 * exp(-0.02 * (u + 30)) / (1 + exp(-0.2 * (u + 30)))
    out["d"] = _lr_gate(st["d"], dt, alpha_d, beta_d)
    out["f"] = _lr_gate(st["f"], dt, alpha_f, beta_f)
    out["cai"] = cai + dt * (-0.0001 * I_Si + 0.07 * (0.0001 - cai))
    # calc_ik
    E_K = R * T / F * log((ko + p["PR_NaK"] * nao) / (ki + p["PR_NaK"] * nai))
    G_K = p["gk"] * sqrt(ko / 5.4)
    if u > -100:
        Xi = 2.837 * (exp(0.04 * (u + 77)) - 1) / ((u + 77) * exp(0.04 * (u + 35)))
    else:
        Xi = 1
    I_K = G_K * st["x"] * Xi * (u - E_K)
    alpha_x = 0.0005 * exp(0.083 * (u + 50)) / (1 + exp(0.057 * (u + 50)))
    beta_x = 0.0013 * exp(-0.06 * (u + 20)) / (1 + exp(-0.04 * (u + 20)))
    out["x"] = _lr_gate(st["x"], dt, alpha_x, beta_x)
    E_K1 = R * T / F * log(ko / ki)
    # calc_ik1, calc_ikp, calc_ib
    alpha_K1 = 1.02 / (1 + exp(0.2385 * (u - E_K1 - 59.215)))
    beta_K1 = (0.49124 * exp(0.08032 * (u - E_K1 + 5.476)) +
               exp(0.06175 * (u - E_K1 - 594.31))) / (1 + exp(-0.5143 * (u - E_K1 + 4.753)))
    K_1x = alpha_K1 / (alpha_K1 + beta_K1)
    ik1 = p["gk1"] * sqrt(ko / 5.4) * K_1x * (u - E_K1)
    ikp = p["gkp"] * (1.0 / (1 + exp((7.488 - u) / 5.98))) * (u - E_K1)
    ib = p["gb"] * (u + 59.87)
    ik1t = ik1 + ikp + ib
    return -(dt * (ina + I_Si + ik1t + I_K)), out


# ---------------------------------------------------------------- Courtemanche
def _crn_gate(x, x_inf, tau_x, dt):
    # courtemanche_2d.py calc_gating_variable: Rush-Larsen
    return x_inf - (x_inf - x) * math.exp(-dt / tau_x)


def courtemanche(u, st, dt, p):
    """courtemanche_2d.py:180-641.  st: the kernel's slots nai, ki, cai,
    caup, carel, m, h, j_, d, f, oa, oi, ua, ui, xs, xr, fca, irel, vrel,
    urel, wrel.  calc_ikr ignores its gkr argument and uses 0.0294 (:401)."""
    exp, log, sqrt = math.exp, math.log, math.sqrt
    R, T, F = p["R"], p["T"], p["F"]
    nao, ko, cao = p["nao"], p["ko"], p["cao"]
    kq10 = p["kq10"]
    out = dict(st)
    nai, ki, cai, caup, carel = st["nai"], st["ki"], st["cai"], st["caup"], st["carel"]
    ena = R * T / F * log(nao / nai)
    ek = R * T / F * log(ko / ki)
    eca = R * T / (2 * F) * log(cao / cai)
    # m, h, j
    if u == -47.13:
        am = 3.2
    else:
        am = 0.32 * (u + 47.13) / (1 - exp(-0.1 * (u + 47.13)))
    bm = 0.08 * exp(-u / 11)
    m = _crn_gate(st["m"], am / (am + bm), 1 / (am + bm), dt)
    if u >= -40:
        ah = 0
        bh = 1 / (0.13 * (1 + exp(-(u + 10.66) / 11.1)))
    else:
        ah = 0.135 * exp(-(80 + u) / 6.8)
        bh = 3.56 * exp(0.079 * u) + 310000 * exp(0.35 * u)
    h = _crn_gate(st["h"], ah / (ah + bh), 1 / (ah + bh), dt)
    if u >= -40:
        aj = 0
        bj = 0.3 * exp(-2.535e-07 * u) / (1 + exp(-0.1 * (u + 32)))
    else:
        aj = (-127140 * exp(0.2444 * u) - 3.474e-05 * exp(-0.04391 * u)) * (u + 37.78) / \
            (1 + exp(0.311 * (u + 79.23)))
        bj = 0.1212 * exp(-0.01052 * u) / (1 + exp(-0.1378 * (u + 40.14)))
    j_inf = aj / (aj + bj)
    tau_j = 1 / (aj + bj)
    j = j_inf - (j_inf - st["j_"]) * exp(-dt / tau_j)
    out.update(m=m, h=h, j_=j)
    ina = p["gna"] * m ** 3 * h * j * (u - ena)
    ik1 = p["gk1"] * (u - ek) / (1 + exp(0.07 * (u + 80)))
    # ito
    ao = 0.65 / (exp(-(u + 10) / 8.5) + exp(-(u - 30) / 59.0))
    bo = 0.65 / (2.5 + exp((u + 82) / 17.0))
    tau_o = 1 / (kq10 * (ao + bo))
    o_inf = 1 / (1 + exp(-(u + 20.47) / 17.54))
    aoi = 1 / (18.53 + exp((u + 113.7) / 10.95))
    boi = 1 / (35.56 + exp(-(u + 1.26) / 7.44))
    tau_oi = 1 / (kq10 * (aoi + boi))
    oi_inf = 1 / (1 + exp((u + 43.1) / 5.3))
    oa = _crn_gate(st["oa"], o_inf, tau_o, dt)
    oi = _crn_gate(st["oi"], oi_inf, tau_oi, dt)
    ito = p["gto"] * oa ** 3 * oi * (u - ek)
    out.update(oa=oa, oi=oi)
    # ikur
    gkur = 0.005 + 0.05 / (1 + exp(-(u - 15) / 13.0))
    aua = 0.65 / (exp(-(u + 10) / 8.5) + exp(-(u - 30) / 59.0))
    bua = 0.65 / (2.5 + exp((u + 82) / 17.0))
    tau_ua = 1 / (kq10 * (aua + bua))
    ua_inf = 1 / (1 + exp(-(u + 30.3) / 9.6))
    aui = 1 / (21 + exp(-(u - 185) / 28.0))
    bui = exp((u - 158) / 16.0)
    tau_ui = 1 / (kq10 * (aui + bui))
    ui_inf = 1 / (1 + exp((u - 99.45) / 27.48))
    ua = _crn_gate(st["ua"], ua_inf, tau_ua, dt)
    ui = _crn_gate(st["ui"], ui_inf, tau_ui, dt)
    ikur = p["gkur_coeff"] * gkur * ua ** 3 * ui * (u - ek)
    out.update(ua=ua, ui=ui)
    # ikr (slot xr), gkr fixed at 0.0294
    axr = 0.0003 * (u + 14.1) / (1 - exp(-(u + 14.1) / 5))
    bxr = 7.3898e-05 * (u - 3.3328) / (exp((u - 3.3328) / 5.1237) - 1)
    xr = _crn_gate(st["xr"], 1 / (1 + exp(-(u + 14.1) / 6.5)), 1 / (axr + bxr), dt)
    ikr = 0.0294 * xr * (u - ek) / (1 + exp((u + 15) / 22.4))
    # iks (slot xs)
    axs = 4e-05 * (u - 19.9) / (1 - exp(-(u - 19.9) / 17))
    bxs = 3.5e-05 * (u - 19.9) / (exp((u - 19.9) / 9) - 1)
    xs = _crn_gate(st["xs"], 1 / sqrt(1 + exp(-(u - 19.9) / 12.7)), 1 / (2 * (axs + bxs)), dt)
    iks = p["gks"] * xs ** 2 * (u - ek)
    out.update(xr=xr, xs=xs)
    # ical
    tau_d = (1 - exp(-(u + 10) / 6.24)) / (0.035 * (u + 10) * (1 + exp(-(u + 10) / 6.24)))
    d_inf = 1 / (1 + exp(-(u + 10) / 8.0))
    tau_f = 9 / (0.0197 * exp(-0.0337 ** 2 * (u + 10) ** 2) + 0.02)
    f_inf = 1 / (1 + exp((u + 28) / 6.9))
    fca_inf = 1 / (1 + cai / 0.00035)
    d = _crn_gate(st["d"], d_inf, tau_d, dt)
    f = _crn_gate(st["f"], f_inf, tau_f, dt)
    fca = _crn_gate(st["fca"], fca_inf, 2, dt)
    ical = p["gcal"] * d * f * fca * (u - 65)
    out.update(d=d, f=f, fca=fca)
    # inak
    s = 1 / 7.0 * (exp(nao / 67.3) - 1)
    fnak = 1 / (1 + 0.1245 * exp(-0.1 * (F * u) / (R * T)) + 0.0365 * s * exp(-(F * u) / (R * T)))
    inak = p["inakmax"] * fnak * (1 / (1 + (p["kmnai"] / nai) ** 1.5)) * (ko / (ko + p["kmko"]))
    # inaca
    gamma = 0.35
    exp_term = exp(gamma * (F * u) / (R * T))
    exp_rev_term = exp((gamma - 1) * (F * u) / (R * T))
    numerator = p["inacamax"] * (exp_term * nai ** 3 * cao - exp_rev_term * nao ** 3 * cai)
    denominator = (p["kmnancx"] ** 3 + nao ** 3) * (p["kmcancx"] + cao) * \
        (1 + p["ksatncx"] * exp_rev_term)
    inaca = numerator / denominator
    ibca = p["gcab"] * (u - eca)
    ibna = p["gnab"] * (u - ena)
    ipca = p["ipcamax"] * cai / (cai + 0.0005)
    # irel
    Vrel = p["Vrel"]
    Fn = 1e-12 * Vrel * st["irel"] - 5 * 1e-13 / F * (0.5 * ical - 0.2 * inaca)
    u_inf = 1 / (1 + exp(-(Fn - 3.4175e-13) / 1.367e-15))
    tau_v = 1.91 + 2.09 / (1 + exp(-(Fn - 3.4175e-13) / 1.367e-15))
    v_inf = 1 - 1 / (1 + exp(-(Fn - 6.835e-14) / 1.367e-15))
    tau_w = 6 * (1 - exp(-(u - 7.9) / 5.0)) / ((1 + 0.3 * exp(-(u - 7.9) / 5.0)) * (u - 7.9))
    w_inf = 1 - 1 / (1 + exp(-(u - 40) / 17.0))
    urel = _crn_gate(st["urel"], u_inf, 8, dt)
    vrel = _crn_gate(st["vrel"], v_inf, tau_v, dt)
    wrel = _crn_gate(st["wrel"], w_inf, tau_w, dt)
    irel = p["krel"] * urel ** 2 * vrel * wrel * (carel - cai)
    out.update(irel=irel, urel=urel, vrel=vrel, wrel=wrel)
    itr = (caup - carel) / 180
    iupmax, Vup, Vj = p["iupmax"], p["Vup"], p["Vj"]
    iup = iupmax / (1 + p["kup"] / cai)
    iupleak = caup / p["caupmax"] * iupmax
    out["caup"] = caup + dt * (iup - iupleak - itr * (Vrel / Vup))
    out["nai"] = nai + dt * ((-3 * inak - 3 * inaca - ibna - ina) / (F * Vj))
    out["ki"] = ki + dt * ((2 * inak - ik1 - ito - ikur - ikr - iks - p["ibk"]) / (F * Vj))
    B1 = (2 * inaca - ipca - ical - ibca) / (2 * F * Vj) + (Vup * (iupleak - iup) + irel * Vrel) / Vj
    B2 = 1 + p["trpnmax"] * p["kmtrpn"] / (cai + p["kmtrpn"]) ** 2 + \
        p["cmdnmax"] * p["kmcmdn"] / (cai + p["kmcmdn"]) ** 2
    out["cai"] = cai + dt * (B1 / B2)
    out["carel"] = carel + dt * ((itr - irel) / (1 + p["csqnmax"] * p["kmcsqn"] /
                                                 (carel + p["kmcsqn"]) ** 2))
    du = -(dt * (ina + ik1 + ito + ikur + ikr + iks + ical + ipca + inak + inaca + ibna + ibca))
    return du, out
