"""Known answers for the control layer of GmresLDLIR::Solve / PrecondGmres
(include/MPIR/gmres_ir.hpp:534-741) and Compute's scaling (:347-363).

Each trace below was worked out by hand from the reference's text, not from
any implementation; a plausible misreading of a rule changes at least one
number of some case (the comment of each case names the misreadings it
separates). The oracle tests and the GPU tests both assert them.

Stop reasons: 0 iteration limit, 1 small update, 2 residual stagnation,
3 exactly zero residual (a guard of this implementation; the reference
would divide by the zero norm and continue with NaN).
"""
import numpy as np

FLT_EPS = float(np.finfo(np.float32).eps)

# K = I (n = 4, upper CSC), b = 0.5 in every entry: ||b||_2 = 1, ||b||_inf = 0.5.
# The scaling leaves K alone (a_ii = 1), M^{-1} = I exactly, and every
# Krylov quantity is exact in binary: V0 = b / ||b||_2, h00 = 1, w = 0.
IDENTITY_N = 4
IDENTITY_B = 0.5

# tol -> (x equals b (else 0), IR iterations, GMRES iterations, M^{-1} applications, stop reason)
IDENTITY_TRACES = {
    # ||b||_2 = 1 >= 0.75: the first PrecondGmres runs one Arnoldi step
    # (happy breakdown) and returns b exactly; the residual is exactly zero.
    # An InfNrm early exit (0.5 < 0.75) would return 0 first: IR 2.
    0.75: (True, 1, 1, 2, 3),
    # ||b||_2 = 1 is not < 1: no early exit (a `<=` rule exits: 2 applications,
    # not 3); V0 is formed (1 application), but rho = 1 > 1 fails, so no
    # Arnoldi step: x0 = 0. IR 1: b' = r / 0.5 has ||b'||_2 = 2, one step
    # gives x = b exactly; IR 2 sees a zero residual.
    1.0: (True, 2, 1, 3, 3),
    # early exit on the first call (x0 = 0, no application); IR 1 runs one
    # step on b' (||b'||_2 = 2 >= 1.5); IR 2 sees a zero residual.
    1.5: (True, 2, 1, 2, 3),
    # every call exits early: x0 = 0, d = 0, and d_norm = 0 <= 10 eps * 0
    # stops IR 1 with reason 1 (a strict `<` would go on to IR 2 and stop
    # there on |r_norm - r_prev| = 0 < tol, reason 2).
    2.5: (False, 1, 0, 0, 1),
}


def identity_upper(n):
    return np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.ones(n)


# Compute's scaling threshold |a_ii| > eps(UF), strict, on |a_ii|:
# diagonal values and their expected AD per factor precision.
SCALING_DIAG = np.array([4.0, 1e-8, -1e-8, FLT_EPS, 2 * FLT_EPS, 2.0 ** -10, 2.0 ** -9, 1e-12])
SCALING_AD = {
    # eps(double): every nonzero diagonal here scales
    np.float64: [0.5, 1e4, 1e4, FLT_EPS ** -0.5, (2 * FLT_EPS) ** -0.5, 32.0, 2.0 ** 4.5, 1e6],
    # eps(float) = 2^-23: 1e-8 stays unscaled (an eps(UW) threshold scales it),
    # and so does a_ii = eps exactly (a `>=` threshold scales it)
    np.float32: [0.5, 1.0, 1.0, 1.0, (2 * FLT_EPS) ** -0.5, 32.0, 2.0 ** 4.5, 1.0],
    # eps(binary16) = 2^-10
    np.float16: [0.5, 1.0, 1.0, 1.0, 1.0, 1.0, 2.0 ** 4.5, 1.0],
    # eps(bfloat16) = 2^-7
    "bf16": [0.5, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0],
}
# one off-diagonal entry a_01 = 3 couples rows 0 and 1: scaled 3 * AD0 * AD1
SCALING_A01 = 3.0


def scaling_matrix():
    """Upper CSC: the diagonal SCALING_DIAG plus a_01 = 3."""
    n = len(SCALING_DIAG)
    Ap = [0]
    Ai, Ax = [], []
    for j in range(n):
        if j == 1:
            Ai.append(0)
            Ax.append(SCALING_A01)
        Ai.append(j)
        Ax.append(SCALING_DIAG[j])
        Ap.append(len(Ai))
    return np.array(Ap, np.int32), np.array(Ai, np.int32), np.array(Ax)


# Absolute stagnation stop |r_norm - r_prev| < tol (gmres_ir.hpp:606) on the
# n = 1280 KKT system with an FP32 factor and tol = 1e-3: after IR 1 the
# relative residual is already far below tol (so an `r_norm < tol` rule would
# stop at IR 1), and r_prev starts at DBL_MAX (so IR 1 cannot stagnate); the
# reference stops at IR 2 with reason 2. A relative stagnation test
# |r - r_prev| / r_prev < tol would not stop at IR 2 either.
STAGNATION_TOL = 1e-3
STAGNATION_IR = 2
STAGNATION_REASON = 2
