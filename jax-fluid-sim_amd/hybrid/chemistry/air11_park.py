"""Park 1990 11-species, 22-reaction air mechanism data
(hybrid/chemistry/air11_park.py:11-311 of the reference: the same species
order, NASA-7 fits, Arrhenius constants, third-body efficiencies and
temperature modes), packed into the table layout the chemistry kernels read.
"""

import numpy as np

R_UNIV = 8.31446261815324  # J / (mol K), the kernels' gas constant

MW_E = 5.48579909065e-7
MW_N = 0.0140067
MW_O = 0.015999
MW_NO = MW_N + MW_O
MW_N2 = 2.0 * MW_N
MW_O2 = 2.0 * MW_O

SPECIES_NAMES = ("e-", "N+", "O+", "NO+", "N2+", "O2+", "N", "O", "NO", "N2", "O2")
SP_E, SP_NP, SP_OP, SP_NOP, SP_N2P, SP_O2P, SP_N, SP_O, SP_NO, SP_N2, SP_O2 = range(11)
MOLAR_MASSES = (MW_E, MW_N, MW_O, MW_NO, MW_N2, MW_O2, MW_N, MW_O, MW_NO, MW_N2, MW_O2)
CHARGES = (-1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0)

_NASA = (
    (298.15, 1000.0, 6000.0, (2.5, 0.0, 0.0, 0.0, 0.0, -745.375, -11.7208122),
     (2.5, 0.0, 0.0, 0.0, 0.0, -745.375, -11.7208122)),
    (298.15, 1000.0, 6000.0,
     (2.79781357, -1.41287178e-03, 2.68686998e-06, -2.31615591e-09, 7.49585578e-13, 225548.731, 3.59928845),
     (2.51210704, 1.75367440e-06, -1.50557538e-08, 7.06834189e-12, -6.31861166e-16, 225597.004, 4.92236293)),
    (298.15, 1000.0, 6000.0, (2.5, 0.0, 0.0, 0.0, 0.0, 187935.284, 4.39337676),
     (2.48542028, 2.56978695e-05, -1.28833378e-08, 1.65525487e-12, 1.09933344e-16, 187940.874, 4.47425446)),
    (298.15, 1000.0, 6000.0,
     (3.69301231, -1.34229158e-03, 2.67343395e-06, -1.02609308e-09, -6.95610492e-14, 118103.055, 3.09126691),
     (2.94587702, 1.40325260e-03, -4.95503196e-07, 7.95948973e-11, -4.72076668e-15, 118244.340, 6.70644634)),
    (298.15, 1000.0, 6000.0,
     (3.77540711, -2.06459157e-03, 4.75752301e-06, -3.15664228e-09, 6.70509973e-13, 180481.115, 2.69322178),
     (3.58661363, 2.53071949e-04, 1.84778214e-07, -4.55257223e-11, 3.26818029e-15, 180390.994, 3.09584143)),
    (298.15, 1000.0, 6000.0,
     (4.61017167, -6.35951952e-03, 1.42425624e-05, -1.20997923e-08, 3.70956878e-12, 139742.229, -0.201326941),
     (3.31675922, 1.11522244e-03, -3.83492556e-07, 5.72784687e-11, -2.77648381e-15, 139876.823, 5.44726469)),
    (200.0, 1000.0, 6000.0, (2.5, 0.0, 0.0, 0.0, 0.0, 56104.638, 4.1939088),
     (2.4159429, 1.74890650e-04, -1.19023690e-07, 3.02262440e-11, -2.03609830e-15, 56133.775, 4.6496095)),
    (200.0, 1000.0, 6000.0,
     (3.16826710, -3.27931884e-03, 6.64306396e-06, -6.12806624e-09, 2.11265971e-12, 29122.2592, 2.05193346),
     (2.54363697, -2.73162486e-05, -4.19029520e-09, 4.95481845e-12, -4.79553694e-16, 29226.012, 4.92229457)),
    (200.0, 1000.0, 6000.0,
     (4.21859896, -4.63988124e-03, 1.10443049e-05, -9.34055507e-09, 2.80554874e-12, 9845.09964, 2.28061001),
     (3.26071234, 1.19101135e-03, -4.29122646e-07, 6.94481463e-11, -4.03295681e-15, 9921.43132, 6.36900518)),
    (200.0, 1000.0, 6000.0,
     (3.53100528, -1.23660988e-04, -5.02999433e-07, 2.43530612e-09, -1.40881235e-12, -1046.97628, 2.96747038),
     (2.95257637, 1.39690040e-03, -4.92631603e-07, 7.86010195e-11, -4.60755204e-15, -923.948688, 5.87188762)),
    (200.0, 1000.0, 6000.0,
     (3.78245636, -2.99673416e-03, 9.84730201e-06, -9.68129509e-09, 3.24372837e-12, -1063.94356, 3.65767573),
     (3.66096065, 6.56365811e-04, -1.41149627e-07, 2.05797935e-11, -1.29913436e-15, -1215.97718, 3.41536279)),
)

# (reactants, products, A, b, theta, third body, efficiencies, fwd mode, rev mode);
# modes 0 = T, 1 = sqrt(T Tv), 2 = Tv.
_E, _NP, _OP, _NOP, _N2P, _O2P, _N, _O, _NO, _N2, _O2 = range(11)
_REACTIONS = (
    ({_N2: 1}, {_N: 2}, 7.0e15, -1.6, 113200.0, 1, {_O: 4.285714285714286, _N: 4.285714285714286}, 1, 0),
    ({_N2: 1, _E: 1}, {_N: 2, _E: 1}, 3.0e18, -1.6, 113200.0, 0, {}, 2, 2),
    ({_O2: 1}, {_O: 2}, 2.0e15, -1.5, 59360.0, 1, {_O: 5.0, _N: 5.0}, 1, 0),
    ({_NO: 1}, {_N: 1, _O: 1}, 5.0e9, 0.0, 75500.0, 1,
     {_NO: 22.0, _N: 22.0, _O: 22.0, _NP: 22.0, _OP: 22.0}, 1, 0),
    ({_N2: 1, _O: 1}, {_NO: 1, _N: 1}, 5.7e6, 0.42, 42938.0, 0, {}, 0, 0),
    ({_NO: 1, _O: 1}, {_O2: 1, _N: 1}, 8.4e6, 0.0, 19400.0, 0, {}, 0, 0),
    ({_N: 1, _O: 1}, {_NOP: 1, _E: 1}, 5.3e6, 0.0, 31900.0, 0, {}, 0, 2),
    ({_O: 2}, {_O2P: 1, _E: 1}, 7.1e-4, 2.7, 80600.0, 0, {}, 0, 2),
    ({_N: 2}, {_N2P: 1, _E: 1}, 4.4e1, 1.5, 67500.0, 0, {}, 0, 2),
    ({_NOP: 1, _O: 1}, {_NP: 1, _O2: 1}, 1.0e6, 0.5, 77200.0, 0, {}, 0, 0),
    ({_NP: 1, _N2: 1}, {_N2P: 1, _N: 1}, 1.0e6, 0.5, 12200.0, 0, {}, 0, 0),
    ({_O2P: 1, _N: 1}, {_NP: 1, _O2: 1}, 8.7e7, 0.14, 28600.0, 0, {}, 0, 0),
    ({_OP: 1, _NO: 1}, {_NP: 1, _O2: 1}, 1.4e-1, 1.9, 26600.0, 0, {}, 0, 0),
    ({_O2P: 1, _N2: 1}, {_N2P: 1, _O2: 1}, 9.9e6, 0.0, 40700.0, 0, {}, 0, 0),
    ({_O2P: 1, _O: 1}, {_OP: 1, _O2: 1}, 4.0e6, 0.09, 18000.0, 0, {}, 0, 0),
    ({_NOP: 1, _N: 1}, {_OP: 1, _N2: 1}, 3.4e7, -1.08, 12800.0, 0, {}, 0, 0),
    ({_NOP: 1, _O2: 1}, {_O2P: 1, _NO: 1}, 2.4e7, 0.41, 32600.0, 0, {}, 0, 0),
    ({_NOP: 1, _O: 1}, {_O2P: 1, _N: 1}, 7.2e6, 0.29, 48600.0, 0, {}, 0, 0),
    ({_OP: 1, _N2: 1}, {_N2P: 1, _O: 1}, 9.1e5, 0.36, 22800.0, 0, {}, 0, 0),
    ({_NOP: 1, _N: 1}, {_N2P: 1, _O: 1}, 7.2e7, 0.0, 35500.0, 0, {}, 0, 0),
    ({_O: 1, _E: 1}, {_OP: 1, _E: 2}, 3.9e27, -3.78, 158500.0, 0, {}, 2, 0),
    ({_N: 1, _E: 1}, {_NP: 1, _E: 2}, 2.5e28, -3.82, 168200.0, 0, {}, 2, 0),
)


def stoichiometry():
    """(nu', nu'', third-body efficiencies, third-body flags) as (22, 11)
    float64 arrays and a bool vector."""
    ns, nr = 11, len(_REACTIONS)
    nu1, nu2, alpha = np.zeros((nr, ns)), np.zeros((nr, ns)), np.zeros((nr, ns))
    third = np.zeros(nr, bool)
    for r, (rea, pro, A, b, th, t3, eff, fm, rm) in enumerate(_REACTIONS):
        for s, c in rea.items():
            nu1[r, s] += c
        for s, c in pro.items():
            nu2[r, s] += c
        third[r] = bool(t3)
        if t3:
            alpha[r, 1:] = 1.0          # every species but the electron
            for s, e in eff.items():
                alpha[r, s] = e
    return nu1, nu2, alpha, third


def stoichiometry_fingerprint(nu1, nu2, alpha, third):
    """One float64 that identifies the stoichiometry and third-body data: the
    kernels that compile these in (csrc/fsg_park11.h) compare it with the
    table's last entry.  Summed in a fixed order, so it is reproducible."""
    nr, ns = nu1.shape
    fp = 0.0
    for r in range(nr):
        fp += 1000.0 * (r + 1) * float(bool(third[r]))
        for s in range(ns):
            fp += (r * ns + s + 1) * (float(nu1[r, s]) + 2.0 * float(nu2[r, s])
                                      + 4.0 * float(alpha[r, s]))
    return fp


def mechanism_table():
    """Flat float64 table in the layout of csrc/fsg_chem.hip."""
    ns, nr = 11, len(_REACTIONS)
    nasa = []
    for lo, mid, hi, cl, ch in _NASA:
        nasa += [lo, mid, hi, *cl, *ch]
    rx = []
    for r, (rea, pro, A, b, th, third, eff, fm, rm) in enumerate(_REACTIONS):
        rx += [A, b, th, float(third), float(fm), float(rm)]
    nu1, nu2, alpha, third_flags = stoichiometry()
    # derived per reaction: ln A, sum(nu'' - nu'), sum M (nu'' - nu')
    d = nu2 - nu1
    derived = np.stack([np.log([r[2] for r in _REACTIONS]), d.sum(1),
                        d @ np.asarray(MOLAR_MASSES)], 1)
    # derived per species: 1/M, R/M, 1/t_low, 1/t_high, then per fit range
    # (low, high) the divided NASA-7 coefficients of h/RT (a1/2, a2/3, a3/4,
    # a4/5) and of s/R (a2/2, a3/3, a4/4) for the kernels' Horner forms
    spec = []
    for M, (lo, mid, hi, cl, ch) in zip(MOLAR_MASSES, _NASA):
        spec += [1.0 / M, R_UNIV / M, 1.0 / lo, 1.0 / hi]
        for a in (cl, ch):
            spec += [a[1] / 2.0, a[2] / 3.0, a[3] / 4.0, a[4] / 5.0, a[2] / 2.0, a[3] / 3.0,
                     a[4] / 4.0]
    # last: the stoichiometry fingerprint (csrc/fsg_park11.h)
    fp = stoichiometry_fingerprint(nu1, nu2, alpha, third_flags)
    return np.concatenate([np.asarray(MOLAR_MASSES), nasa, rx, nu1.ravel(), nu2.ravel(),
                           alpha.ravel(), derived.ravel(), spec, [fp]]).astype(np.float64)
