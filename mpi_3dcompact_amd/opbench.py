"""Per-operator throughput of the derivative / filter line operators.

    python -m mpi_3dcompact_amd.opbench [--n 256] [--fields 5] [--reps 20]
                                        [--only penta10] [--bc periodic]
    python -m mpi_3dcompact_amd.opbench --kernel-level [--n 256]

For every scheme, direction and boundary type it applies one operator to
``--fields`` 256^3 FP64 fields (one batched launch) and prints the mean time
per field and the algorithmic bandwidth (16 B per point per field: the input
read once, the output written once).  Plain device-to-device copies of one
field and of ``--fields`` fields are timed first: the many-field copy is the
practical bandwidth reference for a launch of that many fields, and the
one-field copy's hardware counters (rocprofv3 FETCH_SIZE / WRITE_SIZE)
calibrate those of the operators.

--kernel-level times the kernel-level workload as one pass: Penta10 first
derivatives in xi, eta and zeta and the Compact8 filter (alpha_f 0.49) in all
three directions, each on five fields, periodic in xi and with the boundary
closures in eta and zeta (six launches), and prints the pass time and its
aggregate algorithmic bandwidth (6 x 16 B per point per field).
"""
import argparse
import sys

SCHEMES = {  # name -> (class, kind, alpha_f)
    "penta10": ("deriv", 2, 0.0),
    "pade6": ("deriv", 1, 0.0),
    "cd2": ("deriv", 0, 0.0),
    "compact8": ("filter", 0, 0.49),
    "compact10": ("filter", 1, 0.49),
    "commn5pade": ("filter", 2, 0.0),
    "commn5exp": ("filter", 3, 0.0),
}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--fields", type=int, default=5)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--only", default="", help="comma-separated scheme names")
    ap.add_argument("--bc", default="periodic,dirichlet")
    ap.add_argument("--kernel-level", action="store_true", help="the six-launch kernel-level pass only")
    args = ap.parse_args(argv)

    import torch
    from .ops import LineOp

    if not torch.cuda.is_available():
        raise SystemExit("opbench needs a GPU")
    N, nf = args.n, args.fields
    dims = (N, N, N)
    npts = N ** 3
    fs = [torch.rand((N, N, N), dtype=torch.float64, device="cuda") for _ in range(nf)]
    outs = [torch.empty_like(f) for f in fs]

    def timeit(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.reps

    if args.kernel_level:
        bc3 = (0, 1, 1)  # periodic xi, closures in eta and zeta
        ops = [LineOp.deriv(2, d, bc3[d], dims, 1.0 / N if bc3[d] == 0 else 1.0 / (N - 1)) for d in range(3)]
        ops += [LineOp.filter(0, 0.49, d, bc3, dims) for d in range(3)]
        derivs = [[torch.empty_like(f) for f in fs] for _ in range(3)]
        filt = [torch.empty_like(f) for f in fs]

        def one_pass():
            for d in range(3):
                ops[d].apply(fs, derivs[d])
            ops[3].apply(fs, filt)
            ops[4].apply(filt, filt)
            ops[5].apply(filt, filt)

        ms = timeit(one_pass)
        for op in ops:
            print(f"  {op.kernel()}", flush=True)
        print(f"kernel-level pass at {N}^3, {nf} fields: {ms:.3f} ms, {6 * 16 * npts * nf / (ms * 1e-3) / 1e9:.0f} GB/s "
              f"algorithmic, {npts / (ms * 1e-3) / 1e6:.0f} Mpt/s", flush=True)
        return 0
    ms = timeit(lambda: outs[0].copy_(fs[0]))
    print(f"copy {N}^3 f64 (1 field): {ms * 1e3:8.1f} us  {16 * npts / (ms * 1e-3) / 1e9:7.0f} GB/s",
          flush=True)
    # the same with as many fields as one operator launch streams: a single
    # 256^3 field pair (268 MB) largely stays in the 256 MB last-level cache
    # across repetitions, so the 1-field copy overstates what HBM delivers
    big_in = torch.rand((nf, N, N, N), dtype=torch.float64, device="cuda")
    big_out = torch.empty_like(big_in)
    ms = timeit(lambda: big_out.copy_(big_in))
    print(f"copy {N}^3 f64 ({nf} fields): {ms * 1e3 / nf:8.1f} us/field  "
          f"{16 * npts * nf / (ms * 1e-3) / 1e9:7.0f} GB/s", flush=True)
    del big_in, big_out
    names = [s for s in args.only.split(",") if s] or list(SCHEMES)
    bcs = [b for b in args.bc.split(",") if b]
    for bname in bcs:
        b = 0 if bname == "periodic" else 1
        bc3 = (b, b, b)
        for name in names:
            cls, kind, af = SCHEMES[name]
            for d in range(3):
                if cls == "deriv":
                    dd = 1.0 / N if b == 0 else 1.0 / (N - 1)
                    op = LineOp.deriv(kind, d, b, dims, dd)
                else:
                    op = LineOp.filter(kind, af, d, bc3, dims)
                ms = timeit(lambda: op.apply(fs, outs))
                gbs = 16 * npts * nf / (ms * 1e-3) / 1e9
                print(f"{bname:9s} {name:10s} dir={d} {op.kernel():34s} {ms * 1e3 / nf:8.1f} us/field "
                      f"{gbs:7.0f} GB/s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
