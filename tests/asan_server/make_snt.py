"""Writes the interaction files the ASan serving driver uses (seeded synthetic
interactions: four on sd 2p2n, three on sd 4p4n, two on pf 2p3n) into the
directory given as the argument."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "nucleartoolkit.jl_amd"))
from smgpu import synth  # noqa: E402

out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "snt")
os.makedirs(out, exist_ok=True)
for k in range(4):
    synth.config_snt(f"{out}/sd_{k}.snt", "sd", 2, 2, 500 + k, (-6.0, 0.0), 1.2)
for k in range(3):
    synth.config_snt(f"{out}/sdb_{k}.snt", "sd", 4, 4, 600 + k, (-6.0, 0.0), 1.2)
for k in range(2):
    synth.config_snt(f"{out}/pf_{k}.snt", "pf", 2, 3, 700 + k, (-6.0, 0.0), 1.2)
