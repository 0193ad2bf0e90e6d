"""Command-line workflow of the reference's example.py: load an OBJ, run
PaMO, write the result (`python -m pamo_amd.example -i in.obj -o out.obj`).

Same options as example.py:13-21.  Without --disable_stage1 the input is
remeshed first (pamo_amd.stage1), so open or non-manifold inputs work; with
it the input must be a closed manifold triangle mesh, as PaMO's stage 2
expects.
"""
from __future__ import annotations

import argparse
import time

import torch

from .obj_io import load_obj, save_obj
from .pamo import PaMO


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("-i", "--input", type=str, required=True)
    p.add_argument("-o", "--output", type=str, required=True)
    p.add_argument("-r", "--ratio", type=float, default=0.1)
    p.add_argument("-mv", "--min-vertex", type=int, default=0)
    p.add_argument("--disable_stage1", action="store_true", help="Disable remeshing")
    p.add_argument("--disable_stage3", action="store_true", help="Disable safe projection")
    p.add_argument("--exact_guard", action="store_true", help="Exact intersection guard")
    args = p.parse_args(argv)
    V, F = load_obj(args.input)
    print(f"# of input verts : {len(V)}")
    print(f"# of input faces : {len(F)}")
    pamo = PaMO((V, F), use_stage1=not args.disable_stage1, use_stage3=not args.disable_stage3,
                exact_guard=args.exact_guard)
    start = time.time()
    verts, faces = pamo.run(torch.from_numpy(V).float().cuda(), torch.from_numpy(F).int().cuda(),
                            min_verts=args.min_vertex, ratio=args.ratio)
    print("Total time: ", time.time() - start)
    print(f"# of output verts : {len(verts)}")
    print(f"# of output faces : {len(faces)}")
    save_obj(args.output, verts, faces)
    return verts, faces


if __name__ == "__main__":
    main()
