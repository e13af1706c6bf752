// GmresLdl with the FGPILU preconditioner: the reference's shipped
// GmresLDLIR::Compute (include/MPIR/gmres_ir.hpp:323-531 of yex33/MPIR):
// the diagonal scaling of Compute, the ILU_k pattern of the matrix in its
// given order, initial guesses from A, then sweeps until the nonlinear
// residual stagnates (|res - res_prev| < tol) or MAX_SWEEPS;
// M^{-1} = U^{-1} L^{-1} (TriangularSolve, :743-774). The GMRES-IR loop
// around it is the one of the L D L^T preconditioner.
//
// In the reference's sweep order every entry reads only values that sweep
// has already made final, so sweep 1 yields the exact ILU factors of the
// pattern and sweep 2 recomputes them bit for bit (its residual equals
// sweep 1's). The GPU runs sweep 1 level by level and reports the sweep
// count the reference's loop reaches: 1 when the first sweep already
// changes the residual by less than tol, else 2 (all MAX_SWEEPS with a
// tolerance of zero, when no change is ever below it).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

#include "host/ilu.hpp"
#include "solver.hpp"

namespace mpir {

namespace {
double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
template <class T>
void upload_vec(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
  b.alloc(std::max<size_t>(1, v.size()) * sizeof(T));
  if (!v.empty()) MPIR_HIP(hipMemcpyAsync(b.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}
}  // namespace

void GmresLdl::set_preconditioner(int kind, int32_t ilu_k, int32_t max_sweeps) {
  MPIR_CHECK(kind == kPrecLdl || kind == kPrecIlu, "preconditioner must be 0 (LDL^T) or 1 (FGPILU)");
  MPIR_CHECK(ilu_k >= 0 && ilu_k <= 8, "ILU level must be 0..8");
  MPIR_CHECK(max_sweeps >= 0, "negative sweep limit");
  MPIR_CHECK(kind == kPrecLdl || uf_ == kF64 || uf_ == kF32, "FGPILU factors are f64 or f32");
  precond_ = kind;
  ilu_k_ = ilu_k;
  ilu_max_sweeps_ = max_sweeps;
  analyzed_ = factored_ = false;
}

void GmresLdl::analyze_ilu(int32_t n, const int32_t* Ap, const int32_t* Ai, double t0) {
  hipStream_t st = stream_;
  IluStructure S;
  try {
    S = ilu_structure(n, Ap, Ai, ilu_k_);
  } catch (const std::exception& e) {
    throw Error(kErrInvalid, e.what());
  }
  // the reference runs FGPILU on the matrix in its given order
  perm_.resize(n);
  std::iota(perm_.begin(), perm_.end(), 0);
  plan_.reset();
  relaxed_ = false;
  Cp_.clear();
  Ci_.clear();
  parent_.clear();
  Lnz_.clear();
  ilu_Up_ = S.Up;
  ilu_Ui_ = S.Ui;
  const std::vector<int32_t>* arrs[9] = {&S.Up, &S.Ui, &S.col, &S.src, &S.ur_ptr, &S.ur_col, &S.ur_idx, &S.lo_rows,
                                         &S.up_rows};
  for (int k = 0; k < 9; ++k) upload_vec(d_ilu_idx_[k], *arrs[k], st);
  ilu_g_ = IluDev();
  ilu_g_.n = n;
  ilu_g_.nnz = S.Up[n];
  ilu_g_.Up = d_ilu_idx_[0].as<int32_t>();
  ilu_g_.Ui = d_ilu_idx_[1].as<int32_t>();
  ilu_g_.col = d_ilu_idx_[2].as<int32_t>();
  ilu_g_.src = d_ilu_idx_[3].as<int32_t>();
  ilu_g_.ur_ptr = d_ilu_idx_[4].as<int32_t>();
  ilu_g_.ur_col = d_ilu_idx_[5].as<int32_t>();
  ilu_g_.ur_idx = d_ilu_idx_[6].as<int32_t>();
  ilu_g_.lo_rows = d_ilu_idx_[7].as<int32_t>();
  ilu_g_.up_rows = d_ilu_idx_[8].as<int32_t>();
  ilu_lv_.h_lo_ptr = S.lo_ptr;
  ilu_lv_.h_up_ptr = S.up_ptr;
  // FP64 CSR of the scaled matrix (identity permutation) for the SpMV
  CscUpper A;
  A.n = n;
  A.p.assign(Ap, Ap + n + 1);
  A.i.assign(Ai, Ai + Ap[n]);
  CsrFull F = csr_from_upper(A);
  upload_vec(d_csr_p_, F.p, st);
  upload_vec(d_csr_j_, F.j, st);
  upload_vec(d_csr_src_, F.src, st);
  d_csr_x_.alloc(std::max<size_t>(1, F.j.size()) * sizeof(double));
  std::vector<int32_t> ecol(nnz_), diag_pos(n, -1);
  for (int32_t j = 0; j < n; ++j)
    for (int32_t q = Ap[j]; q < Ap[j + 1]; ++q) {
      ecol[q] = j;
      if (Ai[q] == j) diag_pos[j] = q;
    }
  upload_vec(d_entry_col_, ecol, st);
  upload_vec(d_entry_row_, Ai_, st);
  upload_vec(d_diag_pos_, diag_pos, st);
  upload_vec(d_perm_, perm_, st);
  // numeric buffers: the shared vectors, A on the pattern, two (L, U, 1/U_jj)
  // sets for the synchronous sweeps
  alloc_vectors();
  const size_t tf = uf_ == kF64 ? sizeof(double) : sizeof(float);
  const int64_t innz = std::max<int64_t>(1, ilu_g_.nnz);
  d_ilu_av_.alloc(innz * tf);
  d_ilu_u_.alloc(innz * tf);
  d_ilu_l_.alloc(innz * tf);
  d_ilu_d_.alloc(n * tf);
  d_ilu_part_.alloc(static_cast<size_t>((innz + 255) / 256 + 1) * sizeof(double));
  d_ilu_out_.alloc(sizeof(double));
  MPIR_HIP(hipStreamSynchronize(st));
  stats_ = SolveStats();
  stats_.n = n;
  stats_.nnz = nnz_;
  stats_.nnz_l = ilu_g_.nnz;
  stats_.nlevels = static_cast<int32_t>(S.lo_ptr.size()) - 1;
  stats_.preconditioner = kPrecIlu;
  stats_.device_bytes = static_cast<int64_t>(3 * innz * tf + n * tf);
  analyzed_ = true;
  precond_only_ = false;
  stats_.analyze_ms = now_ms() - t0;
}

double GmresLdl::ilu_nlres_now(const void* L, const void* U) {
  if (uf_ == kF64)
    ilu_nlres<double>(ilu_g_, d_ilu_av_.as<double>(), static_cast<const double*>(L), static_cast<const double*>(U),
                      d_ilu_part_.as<double>(), d_ilu_out_.as<double>(), stream_);
  else
    ilu_nlres<float>(ilu_g_, d_ilu_av_.as<float>(), static_cast<const float*>(L), static_cast<const float*>(U),
                     d_ilu_part_.as<double>(), d_ilu_out_.as<double>(), stream_);
  MPIR_HIP(hipMemcpyAsync(h_pinned_, d_ilu_out_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
  MPIR_HIP(hipStreamSynchronize(stream_));
  return h_pinned_[0];
}

template <class T>
void GmresLdl::factor_ilu_t() {
  hipStream_t st = stream_;
  T* U = d_ilu_u_.as<T>();
  T* L = d_ilu_l_.as<T>();
  T* Di = d_ilu_d_.as<T>();
  ilu_init<T>(ilu_g_, d_axs_.as<double>(), d_ilu_av_.as<T>(), U, L, Di, st);
  const double res0 = ilu_nlres_now(L, U);
  stats_.ilu_nl_residual0 = res0;
  stats_.ilu_nl_residual = res0;
  stats_.ilu_sweeps = 0;
  if (ilu_max_sweeps_ >= 1) {
    ilu_factor<T>(ilu_g_, ilu_lv_, d_ilu_av_.as<T>(), U, L, Di, st);
    const double res1 = ilu_nlres_now(L, U);
    stats_.ilu_nl_residual = res1;
    if (std::fabs(res1 - res0) < tol_ || ilu_max_sweeps_ == 1) stats_.ilu_sweeps = 1;
    else stats_.ilu_sweeps = 0.0 < tol_ ? 2 : ilu_max_sweeps_;
  }
  // sign count of the pivots U_jj (the signs of 1 / U_jj)
  count_positive<T>(n_, Di, Di, d_cnt_.as<int>(), d_nonfinite_.as<int>(), st);
}

void GmresLdl::factor_ilu() {
  if (uf_ == kF64) factor_ilu_t<double>();
  else factor_ilu_t<float>();
}

void GmresLdl::export_ilu(int32_t* Up, int32_t* Ui, double* Ux, double* Lx, double* UDinv) {
  if (precond_ != kPrecIlu || !factored_) throw Error(kErrState, "no FGPILU factors on this handle");
  MPIR_CHECK(Up && Ui && Ux && Lx && UDinv, "null output");
  const int64_t nnz = ilu_g_.nnz;
  std::memcpy(Up, ilu_Up_.data(), sizeof(int32_t) * (n_ + 1));
  std::memcpy(Ui, ilu_Ui_.data(), sizeof(int32_t) * nnz);
  auto get = [&](const DevBuf& b, int64_t cnt, double* out) {
    if (uf_ == kF64) {
      MPIR_HIP(hipMemcpy(out, b.get(), cnt * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      std::vector<float> tmp(cnt);
      MPIR_HIP(hipMemcpy(tmp.data(), b.get(), cnt * sizeof(float), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < cnt; ++i) out[i] = tmp[i];
    }
  };
  MPIR_HIP(hipStreamSynchronize(stream_));
  get(d_ilu_u_, nnz, Ux);
  get(d_ilu_l_, nnz, Lx);
  get(d_ilu_d_, n_, UDinv);
}

}  // namespace mpir
