"""Writes the converged FP64 oracle images the GPU convergence tests compare with.

    python tests/golden/make_converged_fixtures.py [--threads 8]

Each fixture is the Cornell box (the 256x256 configuration's scene and camera)
rendered by the FP64 oracle at 32x32 with 65,536 samples per pixel, PCG32 seed
1001, depth 8: `cornell_rgb_fp64.npy` (linear RGB) and `cornell_hero_fp64.npy`
(CIE XYZ, hero-wavelength spectral mode). Their Monte Carlo error (relMSE
about 7e-6 against the limit) is far below the 1e-4 bar the tests apply.
"""

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402
import vkrt_amd as V  # noqa: E402
from vkrt_amd import scenes  # noqa: E402

SIZE = 32
SPP = 65536
SEED = 1001
CASES = {
    "cornell_rgb_fp64.npy": dict(mode=V.RENDER_MODE_RGB, spectral=V.SPECTRAL_HERO),
    "cornell_hero_fp64.npy": dict(mode=V.RENDER_MODE_SPECTRAL, spectral=V.SPECTRAL_HERO),
}


def render(case, threads):
    r = V.Renderer(init_device=False)
    scenes.load(r, scenes.cornell_box())
    r.configure(spp=SPP, max_depth=8, min_depth=4, rng=V.RNG_PCG32, seed=SEED, **case)
    r.start(SIZE, SIZE, 0)
    osc = O.OracleScene(r.export(), use_bvh=False)
    acc, bounces = osc.render(SIZE, SIZE, precision=64, threads=threads)
    osc.close()
    r.close()
    return acc[..., :3].astype(np.float32), bounces


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    for name, case in CASES.items():
        t = time.time()
        img, bounces = render(case, args.threads)
        np.save(os.path.join(HERE, name), img)
        print(f"{name}: {bounces} bounces in {time.time() - t:.1f} s, mean {img.reshape(-1, 3).mean(0)}")


if __name__ == "__main__":
    main()
