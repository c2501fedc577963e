"""Command line of the GPU build, mirroring src/diskann.cpp:84-125:

    python -m distributed_diskann_amd [para.json] [--original] [--sweep results.csv]

reads the run parameters (para.json) and the dataset description it names
(dataset.hpp:64-125), loads the base vectors (point_range.h:61-97), builds the
graph on the GPU -- one rank's BSP_index build at P = 1 (bsp_index.hpp:141-434,
with xita, max and start_point_index) by default, knn_index::build_index
(index.h:125-137) with --original, as the reference's two `main` variants --
prints the build time and saves the graph to results/<dataset> in
Graph::save's format (graph.h:178-200).  With --sweep and the dataset's
query and ground-truth files present, it also runs the reference's
search_and_parse protocol and appends its CSV report.  Points are
Euclidian_Point, as in the reference's main.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m distributed_diskann_amd")
    ap.add_argument("para", nargs="?", default="para.json")
    ap.add_argument("--original", action="store_true", help="knn_index::build_index instead of the BSP build")
    ap.add_argument("--sweep", default=None, metavar="CSV", help="run search_and_parse and append its report")
    ap.add_argument("--output", default=None, help="graph file (default results/<dataset>)")
    args = ap.parse_args(argv)

    from . import formats
    from .index import Index
    from .sweep import GraphInfo, search_and_parse

    print(f"Using parameter file: {args.para}", flush=True)
    # dataset.hpp:92 opens data_para relative to the working directory; a
    # path relative to the para file's folder is accepted as well
    dp = formats.DatasetPara().set(args.para)
    ds = dp.dataset_parafile
    if not os.path.isabs(ds) and not os.path.exists(ds):
        ds = os.path.join(os.path.dirname(os.path.abspath(args.para)), ds)
    dp.open(ds)
    print(f"dataset {dp.dataset_name}: nb={dp.nb} nq={dp.nq} d={dp.d} {dp.dtype}; R={dp.R} L={dp.L} "
          f"alpha={dp.alpha} two_pass={dp.two_pass} xita={dp.xita} max={dp.max_build} "
          f"start_point_index={dp.start_point_index}", flush=True)
    X, d = formats.read_vectors(dp.ds_dir, dp.dtype)
    t0 = time.perf_counter()
    idx = Index(X, d, "l2", dp.R)
    if args.original:
        dev_s = idx.build(R=dp.R, L=dp.L, alpha=dp.alpha, two_pass=dp.two_pass)
    else:
        dev_s = idx.build(R=dp.R, L=dp.L, alpha=dp.alpha, start_point=dp.start_point_index, max_fraction=dp.xita,
                          max_build=dp.max_build, final_sort=False, bsp=True)
    wall_ms = (time.perf_counter() - t0) * 1000.0
    print(f"graph build: {wall_ms:.0f} ms ({dev_s * 1000.0:.0f} ms on the device)", flush=True)
    rows = idx.get_graph()
    # build_index_original_parylay always saves (diskann.cpp:69); the BSP
    # build saves when para.json's "save" is set (bsp_index.hpp:472-475)
    if args.original or dp.save:
        out = args.output or dp.output_path
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        formats.write_graph(out, rows)
        print(f"graph saved to {out}", flush=True)

    if args.sweep:
        if not (os.path.exists(dp.qs_dir) and os.path.exists(dp.gt_dir)):
            print("no query / ground-truth file: sweep skipped", file=sys.stderr)
            return 0
        Q, _ = formats.read_vectors(dp.qs_dir, dp.dtype)
        gt_ids, gt_d = formats.read_truth(dp.gt_dir)
        deg = rows[:, 0].astype(np.float64)
        g = GraphInfo(dp.dataset_name, f"R={dp.R} L={dp.L} alpha={dp.alpha}", int(rows.shape[0]), dev_s,
                      float(deg.mean()), int(deg.max()))
        search_and_parse(idx, Q, gt_ids, gt_d.astype(np.float64), k=10, start_point=dp.start_point_index,
                         res_file=args.sweep, graph=g)
    return 0


if __name__ == "__main__":
    sys.exit(main())
