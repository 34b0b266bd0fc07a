"""Residues (reference: pylabolt/utils/residues.py:171-317 and
parallel/cpu/compute_residues_kernels.py:6-73).

Every ``std_out_interval`` steps: ``res = sqrt(sum (phi - phi_old)^2 /
(sum phi_old^2 + float_min))`` per component over fluid nodes, then
``phi_old = phi``.  The sums are FP64 tree reductions on the GPU
(``d2q9_residues``); the reference's order of summation is a parallel
reduction too, so residues agree to rounding, not bit for bit.
"""
import numpy as np
import torch


class ResidueOperator:
    def __init__(self, model, state, backend=None, comm=None, verbose=True):
        self.fields_list = model.residue_fields
        self.residues = {"res_density": np.zeros(1),
                         "res_velocity": np.zeros(2)}
        self.mom_old = None

    def set_backend(self, state, backend, verbose=True):
        # phi_old starts at zero (residues.py:53-64 allocates it with
        # np.zeros_like and set_backend copies it to the device), so the
        # first residue compares against zero fields
        eng = backend.engine
        with torch.cuda.stream(eng.stream):
            self.mom_old = torch.zeros_like(eng.mom[0])

    def compile(self, state, backend, verbose=True):
        return None

    def compute_residues(self, state, backend, mpi_operator, time_step):
        if state.control.std_out_interval is None:
            return
        if time_step % state.control.std_out_interval != 0:
            return
        s = backend.engine.residue_sums(self.mom_old)
        fm = state.control.float_min
        self.residues["res_density"][0] = np.sqrt(s[0] / (s[1] + fm))
        self.residues["res_velocity"][0] = np.sqrt(s[2] / (s[3] + fm))
        self.residues["res_velocity"][1] = np.sqrt(s[4] / (s[5] + fm))
