"""Times the warm-start U-Net's forward pass on the GPU.

    python3 -m artss_amd.unet_bench [--size X Y Z] [--reps N]

Runs the network (random weights, its output layer set non-zero so the
whole graph does real work) on a synthetic two-channel input of X x Y x Z
voxels (each divisible by 8) and prints the mean time of one forward pass
with the nominal operation count of its nine convolutions (2 x 27 x ci x co
per voxel of each layer's level).
"""
import argparse
import os
import time

import numpy as np

import artss_amd

# (ci, co, level) of the nine convolutions, as artss_amd/csrc/unet.cpp lists them
LAYERS = ((2, 16, 0), (16, 16, 0), (16, 32, 1), (32, 64, 2), (64, 64, 3), (64, 64, 2), (64, 32, 1), (32, 16, 0),
          (16, 1, 0))

# any domain gives the network its device and stream
DOMAIN_XML = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs", "taylor_green_64.xml")


def nominal_flops(X, Y, Z):
    n = 0
    for ci, co, level in LAYERS:
        n += 2 * 27 * ci * co * (X >> level) * (Y >> level) * (Z >> level)
    return n


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--size", type=int, nargs=3, default=(256, 128, 128), metavar=("X", "Y", "Z"))
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args()
    X, Y, Z = args.size
    if X % 8 or Y % 8 or Z % 8:
        raise SystemExit("every extent must be divisible by 8")
    dom = artss_amd.Domain(DOMAIN_XML)
    net = artss_amd.UNet(dom, seed=1)
    params = np.random.default_rng(2).standard_normal(net.param_count()) * 0.05
    net.set_params(params)
    x = np.random.default_rng(3).standard_normal((Z, Y, X, 2)).astype(np.float32)
    xb = artss_amd.Buffer(dom, x)
    yb = artss_amd.Buffer(dom, np.zeros((Z, Y, X, 1), np.float32))
    for _ in range(3):
        net.forward(xb, yb)
    dom.sync()
    t0 = time.perf_counter()
    for _ in range(args.reps):
        net.forward(xb, yb)
    dom.sync()
    ms = (time.perf_counter() - t0) / args.reps * 1e3
    gf = nominal_flops(X, Y, Z) / 1e9
    y = yb.download()
    print("forward %dx%dx%d: %.3f ms, %.1f nominal GFLOP, %.1f TFLOP/s nominal, output |y| sum %.6g"
          % (X, Y, Z, ms, gf, gf / ms, float(np.abs(y).sum())), flush=True)


if __name__ == "__main__":
    main()
