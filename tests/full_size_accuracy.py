"""Full-size accuracy of the GPU transforms against the FP64 oracle, one
line of JSON (run from the repository root on a GPU machine:
`python tests/full_size_accuracy.py`).  Not collected by pytest: the
full-size grids take minutes of oracle time; the GPU tests check the same
transforms at smaller sizes and through size-independent properties."""
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


ACCURACY_CASES = [((128, 128, 128), "f64"), ((512, 512, 512), "f32"), ((512, 512, 512), "f64"),
                  ((768, 768, 768), "f32"), ((1024, 512, 2048), "f32"), ((1024, 1024, 1024), "f32")]


def accuracy(cases) -> dict:
    """Relative L2 error of the GPU R2C against the FP64 oracle (pocketfft
    on the same, precision-rounded input) and of the normalised round trip,
    per config grid, at full size."""
    import numpy as np
    import torch
    import gpu_fft_amd as G
    from oracle import gfft_oracle as O
    workers = max(1, min(16, os.cpu_count() or 1))
    out = {}
    for (nx, ny, nz), prec in cases:
        log(f"accuracy {nx}x{ny}x{nz} {prec}")
        rdt = np.float32 if prec == "f32" else np.float64
        fp = O.pad_real(O.random_field(nx, ny, nz, seed=1)).astype(rdt)
        buf = torch.from_numpy(fp).to("cuda:0")
        with G.Plan(nx, ny, nz, precision=prec) as p:
            p.r2c(buf)
            torch.cuda.synchronize()
            got = G.complex_view(buf, nx, ny, nz).cpu().numpy()
            want = O.r2c_p1(fp, nz, workers=workers)
            num = den = 0.0
            for i in range(0, nx, 64):  # chunked norms keep host memory bounded
                d = got[i:i + 64].astype(np.complex128) - want[i:i + 64]
                num += float(np.vdot(d, d).real)
                den += float(np.vdot(want[i:i + 64], want[i:i + 64]).real)
            del got, want
            p.set_scale(1.0, 1.0 / (nx * ny * nz))
            p.c2r(buf)
            torch.cuda.synchronize()
            back = buf.cpu().numpy()[..., :nz]
            rt = O.rel_l2(back, fp[..., :nz])
        out[f"{nx}x{ny}x{nz}_{prec}"] = {"r2c_rel_l2": float(np.sqrt(num / den)), "round_trip_rel_l2": float(rt),
                                         "tolerance": 1e-6 if prec == "f32" else 1e-13}
        del buf, fp, back
        torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    print(json.dumps({"metric": "relative L2 error vs the FP64 oracle per config",
                      "results": accuracy(ACCURACY_CASES)}), flush=True)
