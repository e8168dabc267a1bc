"""Strain of the applied stress, E = S : sigma with the cubic compliance
(cupymag/utils/sigma_matrices.py:25-45)."""
import numpy as np


def get_Ebar_sigma(P):
    S = np.zeros((6, 6))
    S[0, 0] = S[1, 1] = S[2, 2] = P.s11
    S[3, 3] = S[4, 4] = S[5, 5] = P.s44
    S[0, 1] = S[0, 2] = S[1, 0] = S[1, 2] = S[2, 0] = S[2, 1] = P.s12
    return S @ np.array(P.sigma, dtype=np.float64)
