"""Algorithmic HBM bytes of each pass of one RK4 step on the pruned spectrum.

Counts every array a pass must read and write once (the minimum traffic of
that pass), with c = bytes per complex element (16 FP64, 8 FP32):

    N K KZ  fields after the x pass / before the forward x pass
    N N KZ  fields after the y pass / before the forward y pass
    K K KZ  compact state components

pass     reads                      writes
inv_x    3 K K KZ  (6 MHD)          X N K KZ     (X = 5 hydro, 10 MHD)
inv_y    X N K KZ                   F N N KZ     (F = 6 hydro, 12 MHD)
z        F N N KZ                   G N N KZ     (G = 3 hydro, 6 MHD)

The x pass hands over (v, i kx v_y, i kx v_z) per vector v and the y pass
completes i k x v, so X = 5 fields per vector cross the x/y boundary.
fwd_y    G N N KZ                   G N K KZ
fwd_x    G N K KZ + stage vectors   stage vectors

stage vectors (3 K K KZ each, doubled for MHD): stage 0 reads u0, writes
u1, acc (3); stages 1-2 read u1, u0, acc, write u1, acc (5); stage 3 reads
u1, u0, acc, writes u0 (4).
"""

MHD_FACTOR = 2


def sizes(N):
    M = N // 3
    return M, 2 * M + 1, M + 1


def pass_bytes(N, prec_bytes=8, mhd=False):
    """Bytes per launch for each pass name; fwd_x is a list over the 4 stages."""
    _, K, KZ = sizes(N)
    c = 2 * prec_bytes
    X, F, G = (10, 12, 6) if mhd else (5, 6, 3)
    nkk = N * K * KZ * c
    nnk = N * N * KZ * c
    vec = 3 * K * K * KZ * c * (MHD_FACTOR if mhd else 1)
    return {
        "inv_x": vec + X * nkk,
        "inv_y": X * nkk + F * nnk,
        "z": F * nnk + G * nnk,
        "fwd_y": G * nnk + G * nkk,
        "fwd_x": [G * nkk + v * vec for v in (3, 5, 5, 4)],
    }


def step_bytes(N, prec_bytes=8, mhd=False):
    b = pass_bytes(N, prec_bytes, mhd)
    return 4 * (b["inv_x"] + b["inv_y"] + b["z"] + b["fwd_y"]) + sum(b["fwd_x"])


def fft_flops_per_step(N, mhd=False):
    """Nominal 2.5 N^3 log2(N^3) flop per real 3D transform (not pruned), for
    reference only: 9 transforms per hydro RHS, 18 per MHD RHS."""
    import math
    per = 2.5 * N ** 3 * math.log2(N ** 3)
    return 4 * (18 if mhd else 9) * per
