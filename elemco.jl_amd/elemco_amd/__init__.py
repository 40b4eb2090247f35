"""MI355X-native coupled-cluster path of ElemCo.jl: closed-shell CCSD/DCSD, Lambda and (T), and
open-shell UCCSD / UCCD.

The arithmetic lives in the HIP library ``libelemco_hip.so`` (C ABI declared in
``include/elemco_hip.h``); this package is the host-side mirror of the
reference's Julia entry points.
"""
from .lib import (ElemcoError, load, Context, UContext, default_opts, cc_solve_batch, ECG_CCSD, ECG_DCSD, ECG_FP64,
                  ECG_FP32, ECG_BF16X3, ECG_UCCSD, ECG_UCCD, ECG_UDCSD)
from .cc import (ECInfo, ccdriver, ccdriver_unrestricted, calc_MP2, calc_cc, calc_cc_resid, calc_pertT, calc_lm_cc,
                 calc_ccsd_vector_times_Jacobian, calc_1RDM, calc_fock_matrix)
from .fcidump import FDump, read_fcidump, write_fcidump, freeze_orbs_in_dump, parse_orbstring
from .synthetic import make_system
