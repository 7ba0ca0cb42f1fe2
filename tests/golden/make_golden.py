"""Write the golden vectors of tests/golden/*.npz from the reference CPU
backend compiled from its own sources (oracle/Makefile target "ref").

Each fixture holds every input of a forward + adjoint pair and the outputs of
the reference build: receiver traces, final fields and CPML states, saved
histories, adjoint states and the eps/sigma/source gradients.
Run: python tests/golden/make_golden.py   (needs /root/reference)
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import harness as H  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def cases():
    # 2D TM order 2, E-only FP32 history (the configuration of the C1 example)
    p = H.make_problem(24, 22, 2, order=2, nt=30, nshot=1, nsrc=1, nrx=2, seed=101, pml=(4, 4, 4, 4, 0, 0))
    p.storage_type, p.mode = H.RHS_FROM_E, 2
    yield "tm2d_o2_eonly", p
    # 3D order 4, S=3, stored FP32 R, two shots, anisotropic spacing
    p = H.make_problem(14, 13, 12, order=4, nt=10, nshot=2, nsrc=2, nrx=4, seed=102, pml=(3, 2, 2, 3, 2, 3),
                       d=(0.02, 0.017, 0.013))
    p.S = 3
    yield "3d_o4_s3_f32", p
    # 3D order 8, BF16 E+R, PEC cells
    p = H.make_problem(13, 14, 12, order=8, nt=8, nshot=1, nsrc=1, nrx=3, seed=103, pec=True)
    p.storage_type = H.KIND_BF16
    yield "3d_o8_bf16_pec", p
    # 3D order 4 FP16 E+R, zero PML on two faces, x-polarised source, y receivers
    p = H.make_problem(15, 12, 13, order=4, nt=9, nshot=2, nsrc=1, nrx=3, seed=104, pml=(0, 3, 2, 0, 3, 2))
    p.storage_type, p.src_comp, p.rx_comp = H.KIND_F16, 0, 1
    yield "3d_o4_f16_faces", p
    # 2D order 8 general path: coincident sources and receivers, S=2
    p = H.make_problem(22, 19, 2, order=8, nt=12, nshot=2, nsrc=3, nrx=4, seed=105, pml=(3, 3, 3, 3, 0, 0))
    p.src_loc[:, 1] = p.src_loc[:, 0]
    p.rx_loc[:, 2] = p.rx_loc[:, 0]
    p.S = 2
    yield "2d_o8_coincident_s2", p
    # physical-region history, no material gradient (PML forbids it)
    p = H.make_problem(16, 15, 14, order=4, nt=7, nshot=1, nsrc=1, nrx=2, seed=106, pml=(2, 3, 3, 2, 2, 2))
    p.storage_type, p.save_model_history, p.want_eps, p.want_sig = H.PHYSICAL, 0, 0, 0
    yield "3d_o4_physical_fwd", p


def main():
    H.build_oracle(reference=True)
    for name, p in cases():
        out = H.run_cpu(p, "ref")
        path = os.path.join(OUT, f"{name}.npz")
        H.save_case(path, p, out)
        print(name, os.path.getsize(path))


if __name__ == "__main__":
    main()
