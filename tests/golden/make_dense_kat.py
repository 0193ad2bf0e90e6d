"""Writes dense_kat.npz: the quadratic test problem of the reference's only
stage-3 test (simp_cuda/safe_project/tests/test_dummy_energy.py), drawn with
numpy alone in the order that test draws it: Stage3System builds the dummy
energy (A = R^T R with R ~ N(0,1) 6x6, b ~ N(0,1) 6, c ~ N(0,1) 1;
energy.py:83-86) before the test draws x ~ N(0,1) (2,3) as float32.  The
expected answer is the minimiser solve(A, -b) that the test asserts
(atol 1e-6).  Nothing from the reference is imported or run.

    python tests/golden/make_dense_kat.py
"""
import os

import numpy as np

np.random.seed(998244353)
n = 2
A = np.random.randn(n * 3, n * 3)
A = A.T @ A
b = np.random.randn(n * 3)
c = np.random.randn(1)
x0 = np.random.randn(n, 3).astype(np.float32)
x_gt = np.linalg.solve(A, -b).reshape(-1, 3)
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dense_kat.npz")
np.savez(out, A=A, b=b, c=c, x0=x0, x_gt=x_gt)
print(out, np.linalg.cond(A))
