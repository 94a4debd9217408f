"""FP64 restatement of fluidsim's spectral resolution change.

Test infrastructure only (see ``oracle/__init__.py``).

Reference: fill_field_fft_2d (fluidsim/base/init_fields.py:300-314) and the
sequential fill_field_fft_3d (fluidsim/util/mini_oper_modif_resol.py:26-53):
the modes with signed wavenumber indices -m/2 < k <= m/2 along each
signed axis (m = the smaller grid's size) and kx < min(nkx) are copied;
the other modes of the output are left as they are.
"""

import numpy as np


def fill_field_fft_2d(field_in, field_out):
    """init_fields.py:300-314 (loop form)."""
    nk0, nk1 = field_out.shape
    nk0_in, nk1_in = field_in.shape
    nk0_min, nk1_min = min(nk0, nk0_in), min(nk1, nk1_in)
    for ik1 in range(nk1_min):
        field_out[0, ik1] = field_in[0, ik1]
        field_out[nk0_min // 2, ik1] = field_in[nk0_min // 2, ik1]
    for ik0 in range(1, nk0_min // 2):
        for ik1 in range(nk1_min):
            field_out[ik0, ik1] = field_in[ik0, ik1]
            field_out[-ik0, ik1] = field_in[-ik0, ik1]


def fill_field_fft_3d(field_in, field_out):
    """mini_oper_modif_resol.py:26-53 (loop form)."""
    nk0, nk1, nk2 = field_out.shape
    nk0_in, nk1_in, nk2_in = field_in.shape
    m0, m1, m2 = min(nk0, nk0_in), min(nk1, nk1_in), min(nk2, nk2_in)
    for ik0 in range(m0 // 2 + 1):
        for ik1 in range(m1 // 2 + 1):
            for ik2 in range(m2):
                field_out[ik0, ik1, ik2] = field_in[ik0, ik1, ik2]
                if 0 < ik0 < m0 // 2:
                    field_out[-ik0, ik1, ik2] = field_in[-ik0, ik1, ik2]
                    if 0 < ik1 < m1 // 2:
                        field_out[-ik0, -ik1, ik2] = field_in[-ik0, -ik1, ik2]
                if 0 < ik1 < m1 // 2:
                    field_out[ik0, -ik1, ik2] = field_in[ik0, -ik1, ik2]
