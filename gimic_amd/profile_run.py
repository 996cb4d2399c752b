"""Short, fixed GPU workload for hardware-counter runs.

    python -m gimic_amd.profile_run [--points N] [--precision mixed|fp64] [--reps R]

Runs, on cuda:0, (1) a device-to-device copy of exactly 1 GiB (torch), the
reference for scaling FETCH_SIZE / WRITE_SIZE counters to bytes, then (2) R
J-tensor evaluations of the first N points of the 128^3 grid around the
60-atom cage (nbf 1800) in the chosen precision (mixed without the FP64
pass over points next to a nucleus).  With N at most one chunk (on a
288 GB MI355X, the whole grid mixed and 1,048,576 points FP64) every
evaluation is one launch each of the sort, basis evaluation, contraction and
tensor assembly kernels.
"""
import argparse
import time


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=147072)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--kernel", type=int, default=0,
                    help="mixed contraction kernel 3, 4, 5, 8 or 9 (0: automatic selection)")
    ap.add_argument("--split", type=int, default=0, help="N-split count of the mixed kernel")
    a = ap.parse_args()
    import torch
    from . import synth
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print("copy 1 GiB done", flush=True)
    s = synth.cage60()
    ctx = s.context(screening=1e-8)
    pts = torch.tensor(synth.cube_grid(128, 0.12, -7.62)[:a.points], dtype=torch.float64,
                       device="cuda")
    out = torch.empty((pts.shape[0], 9), dtype=torch.float64, device="cuda")
    ctx.set_mixed_kernel(a.kernel, a.split)
    # the counters are for the mixed kernels alone: no FP64 pass near nuclei
    ctx.set_fp64_radius(0.0)
    # first call: kernel selection (mixed) and buffer allocation
    ctx.jtensor_device(pts, precision=a.precision, out=out)
    for i in range(a.reps):
        t0 = time.perf_counter()
        ctx.jtensor_device(pts, precision=a.precision, out=out)
        torch.cuda.synchronize()
        print(f"rep {i}: {pts.shape[0]} points {1e3 * (time.perf_counter() - t0):.1f} ms",
              flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
