"""Benchmark of the learned surrogate: data generation by the plain method
on the GPU (256x256 maps, one particle per cell), training of the periodic
convolutional encoder-decoder on BF16 MFMA kernels, and a rollout on the
1024x1024 map in blocks of K cycles with a plain-method judge.  Prints one
JSON object (losses, TFLOP/s of training and inference against the BF16
MFMA peak, judge results, mass exactness, times).

  python bench_surrogate.py [--terrains 16 --cycles 1300 --K 50 --steps 4000 --blocks 400]
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from simplewinderosion_amd import surrogate  # noqa: E402

if __name__ == "__main__":
    surrogate.main()
