"""Writes the golden fixtures of the Doppler-moment path (u16 windows -> FP64 moments).

Each fixture holds a seeded synthetic window (the speckle generator of oracle/synth.py)
and the moment images of the float64 restatement of the reference graph
(oracle/oracle.py: graph_builder.cc:245-305, 522-566, 578-647), plus the FP64 band
planes after the time DFT.  Run from the repository root:

    python tests/golden/make_golden.py

The files are small (.npz, no pickles) and load with numpy.load(allow_pickle=False).
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from oracle.synth import SynthParams, synth_frames  # noqa: E402

# name -> (synth kwargs, path parameters)
CASES = {
    "fresnel_32x32_t16_shift": (
        dict(height=32, width=32, disk_radius_px=12.0, seed=11, stream_id=0),
        dict(height=32, width=32, window=16, stride=16, bin_begin=3, bin_end=7, prop="fresnel",
             z_m=0.3, fftshift=True)),
    "fresnel_24x48_t32_both_power2": (
        dict(height=24, width=48, disk_radius_px=10.0, seed=12, stream_id=1),
        dict(height=24, width=48, window=32, stride=8, bin_begin=6, bin_end=13, prop="fresnel",
             z_m=0.45, both_sidebands=True, power2=True)),
    "asm_32x32_t32": (
        dict(height=32, width=32, disk_radius_px=12.0, seed=13, stream_id=2),
        dict(height=32, width=32, window=32, stride=32, bin_begin=6, bin_end=13, prop="asm", z_m=0.3)),
    "asm_32x48_pad_filter_split": (
        dict(height=32, width=48, disk_radius_px=12.0, seed=14, stream_id=3),
        dict(height=32, width=48, window=32, stride=8, bin_begin=4, bin_end=14, prop="asm", z_m=0.35,
             pad=(48, 64), filt=(2, 14, 1, 3), split_bin=9, weight_fs_hz=10e3)),
    "asm_16x16_t64_dc_band": (
        dict(height=16, width=16, disk_radius_px=6.0, seed=15, stream_id=4),
        dict(height=16, width=16, window=64, stride=16, bin_begin=0, bin_end=9, prop="asm", z_m=0.2)),
}


def make(name: str):
    sk, pk = CASES[name]
    p = O.PathParams(**pk)
    frames, _ = synth_frames(SynthParams(**sk), 0, p.window)
    planes = O.time_dft(frames, p, "f64").numpy()
    moments = O.doppler_moments(frames, p, "f64")
    return frames, planes, moments, pk


def main():
    index = {}
    for name in CASES:
        frames, planes, moments, pk = make(name)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), window=frames, planes=planes, moments=moments)
        index[name] = {"params": {k: (list(v) if isinstance(v, tuple) else v) for k, v in pk.items()},
                       "synth": CASES[name][0]}
        print(f"{name}: window {frames.shape} moments {moments.shape}")
    with open(os.path.join(HERE, "index.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
