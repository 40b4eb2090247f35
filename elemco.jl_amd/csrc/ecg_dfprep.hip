// AO -> MO transformation of the DF three-index integrals (see ecg_dfprep.h).
#include "ecg_dfprep.h"

#include "ecg_tensor.h"

namespace ecg {

void df_ao_to_mo_device(int device, const double* C, int nao, int nmo, const double* AAP, int nfit,
                        const double* M, int r, double* B) {
  ECG_CHECK(nao > 0 && nmo > 0 && nfit > 0 && r > 0 && r <= nfit, "df_ao_to_mo: invalid sizes");
  ECG_HIP(hipSetDevice(device));
  TensorOps ops;
  ECG_HIP(hipStreamCreate(&ops.stream));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{ops.stream};
  ops.gemm.stream = ops.stream;
  const long aa = (long)nao * nao, mm = (long)nmo * nmo;
  DevBuf dC(sizeof(double) * nao * (long)nmo), dAAP(sizeof(double) * aa * nfit), dM(sizeof(double) * (long)nfit * r);
  DevBuf dAAL(sizeof(double) * aa * r), dmAL(sizeof(double) * (long)nmo * nao * r), dB(sizeof(double) * mm * r);
  ECG_HIP(hipMemcpyAsync(dC.p, C, dC.bytes, hipMemcpyHostToDevice, ops.stream));
  ECG_HIP(hipMemcpyAsync(dAAP.p, AAP, dAAP.bytes, hipMemcpyHostToDevice, ops.stream));
  ECG_HIP(hipMemcpyAsync(dM.p, M, dM.bytes, hipMemcpyHostToDevice, ops.stream));
  // arena for the planner's operand copies: the largest tensor twice
  ops.arena.buf.alloc(sizeof(double) * (2 * std::max(std::max(aa * nfit, aa * r), (long)nmo * nao * r) + 4096));
  Tensor tC(dC.d(), {nao, nmo}), tAAP(dAAP.d(), {nao, nao, nfit}), tM(dM.d(), {nfit, r});
  Tensor tAAL(dAAL.d(), {nao, nao, r}), tmAL(dmAL.d(), {nmo, nao, r}), tB(dB.d(), {nmo, nmo, r});
  // AAL[m,n,L] = AAP[m,n,P] M[P,L]                  (dftools.jl:64)
  ops.contract_p(Prec::F64, 1.0, tAAP, "mnP", tM, "PL", 0.0, tAAL, "mnL");
  ops.arena.reset();
  // mAL[p,n,L] = C[m,p] AAL[m,n,L];  B[p,q,L] = mAL[p,n,L] C[n,q]   (dftools.jl:111-112)
  ops.contract_p(Prec::F64, 1.0, tC, "mp", tAAL, "mnL", 0.0, tmAL, "pnL");
  ops.arena.reset();
  ops.contract_p(Prec::F64, 1.0, tmAL, "pnL", tC, "nq", 0.0, tB, "pqL");
  ECG_HIP(hipMemcpyAsync(B, dB.p, dB.bytes, hipMemcpyDeviceToHost, ops.stream));
  ECG_HIP(hipStreamSynchronize(ops.stream));
}

}  // namespace ecg
