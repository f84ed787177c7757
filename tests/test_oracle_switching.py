"""Pins the oracle's static spin-orbit-torque path (core/junction.hpp:770-812,
static coefficients times the current) on the switching device of the CIMS
ensemble, with an answer the macrospin theory fixes independently of the
code: a perpendicular free layer (anisotropy field H_K = 2K/Ms) under a
damping-like torque and an in-plane bias H_x along the current switches when
the damping-like field exceeds

    H_DL,c = H_K / 2 - |H_x| / sqrt(2)

(K.-S. Lee et al., Appl. Phys. Lett. 102, 112410 (2013), the quasi-static
limit for small damping).  At T = 0 the device must stay at +z at 0.97 of the
corresponding current density and end at -z at 1.03 of it.  Planted mistakes
(a damping-like coefficient off by the TtoAm conversion's 20 %, the bias
sign, the field- and damping-like coefficients swapped) must break it.
Runs on the CPU only."""
from __future__ import annotations

import math

import numpy as np

import oracle
from cmtj_amd import _lib, workloads

TTOAM = 795774.715459
K, MS, HX = 3e5, 1.1, 20e3
J_C = (2 * K / MS / 2 - HX / math.sqrt(2)) / (1e-12 * TTOAM)  # 3.249e11 A/m^2


def _final_mz(plant=None):
    b = workloads.cims_pbit(n_currents=2, n_seeds=1, precision="fp64", j_min=0.97 * J_C, j_max=1.03 * J_C, T=0.0)
    js, layers = b.js.copy(), b.layers.copy()
    if plant:
        plant(js, layers)
    _, mag, st = oracle.run_batch(js, layers, _lib.constants_get(), None, 0, n_threads=2)
    assert (st["code"] == 0).all()
    return mag[:, 0, 2]


def _dl_scaled(js, layers):
    layers["dl_torque"][0::2] *= 0.8


def _bias_flipped(js, layers):
    for s in range(3):
        layers["drv"][:, _lib.SLOT_HEXT + s]["constant"] *= -1


def _fl_dl_swapped(js, layers):
    fl = layers["fl_torque"].copy()
    layers["fl_torque"] = layers["dl_torque"]
    layers["dl_torque"] = fl


def test_sot_switching_threshold_matches_the_macrospin_criterion():
    below, above = _final_mz()
    print(f"j_c = {J_C:.4e} A/m^2: final m_z {below:.4f} at 0.97 j_c, {above:.4f} at 1.03 j_c")
    assert below > 0.9 and above < -0.9
    for p in (_dl_scaled, _bias_flipped, _fl_dl_swapped):
        b2, a2 = _final_mz(p)
        assert not (b2 > 0.9 and a2 < -0.9), p.__name__
