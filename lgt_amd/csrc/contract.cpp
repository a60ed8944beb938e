// Tensor contraction front end: transpose-transpose-GEMM-transpose with a
// layout search that avoids moving the largest operand.
//
// Every pairwise @tensor contraction of the reference becomes one call here.
// The GEMM accepts either operand "k-fast" or "m/n-fast", so a permute is
// only issued when neither storage order of an operand already groups its
// free and contracted labels; among the candidate label orders the one with
// the fewest permuted bytes wins.
#include <algorithm>

#include "peps_internal.h"

namespace peps {

namespace {

struct Plan {
  std::string fa, fb, ct;  // free-A order, free-B order, contracted order
  bool a_kfast, b_kfast;
  bool permA, permB, permC;
  double cost;
};

// Is `lab` a concatenation of x then y (as strings)?
bool is_concat(const std::string& lab, const std::string& x, const std::string& y) {
  return lab == x + y;
}

std::string in_order_of(const std::string& ref, const std::string& set) {
  std::string out;
  for (char ch : ref)
    if (set.find(ch) != std::string::npos) out.push_back(ch);
  return out;
}

std::vector<int> perm_to(const std::string& from, const std::string& to) {
  std::vector<int> p;
  for (char ch : to) p.push_back(int(from.find(ch)));
  return p;
}

}  // namespace

Tensor contract(Ctx& c, const Tensor& A0, const std::string& la0, const Tensor& B0,
                const std::string& lb0, const std::string& lc, bool conjA0, bool conjB0) {
  PEPS_CHECK(int(la0.size()) == A0.rank(), PS_ESHAPE, "contract: labels/rank mismatch for A '" + la0 + "'");
  PEPS_CHECK(int(lb0.size()) == B0.rank(), PS_ESHAPE, "contract: labels/rank mismatch for B '" + lb0 + "'");
  PEPS_CHECK(A0.cplx == B0.cplx, PS_EARG, "contract: mixed real/complex operands");
  // Decide which operand provides the leading output labels.
  bool swap = false;
  if (!lc.empty() && la0.find(lc[0]) == std::string::npos) swap = true;
  const Tensor& A = swap ? B0 : A0;
  const Tensor& B = swap ? A0 : B0;
  const std::string& la = swap ? lb0 : la0;
  const std::string& lb = swap ? la0 : lb0;
  const bool conjA = swap ? conjB0 : conjA0;
  const bool conjB = swap ? conjA0 : conjB0;

  std::string cset, faset, fbset;
  for (char ch : la) {
    if (lb.find(ch) != std::string::npos) cset.push_back(ch);
    else faset.push_back(ch);
  }
  for (char ch : lb)
    if (la.find(ch) == std::string::npos) fbset.push_back(ch);
  PEPS_CHECK(lc.size() == faset.size() + fbset.size(), PS_ESHAPE,
             "contract: output labels '" + lc + "' do not match free labels");
  for (char ch : lc)
    PEPS_CHECK((faset + fbset).find(ch) != std::string::npos, PS_ESHAPE,
               "contract: output label not free");
  std::vector<int64_t> dimof(256, -1);
  for (int i = 0; i < A.rank(); ++i) dimof[(unsigned char)la[i]] = A.dims[i];
  for (int i = 0; i < B.rank(); ++i) {
    auto& d = dimof[(unsigned char)lb[i]];
    PEPS_CHECK(d == -1 || d == B.dims[i], PS_ESHAPE,
               std::string("contract: dimension mismatch on label '") + lb[i] + "' (" + la0 + "," + lb0 + ")");
    d = B.dims[i];
  }
  int64_t M = 1, N = 1, K = 1;
  for (char ch : faset) M *= dimof[(unsigned char)ch];
  for (char ch : fbset) N *= dimof[(unsigned char)ch];
  for (char ch : cset) K *= dimof[(unsigned char)ch];

  // Candidate plans.
  std::vector<std::string> fa_opts = {in_order_of(lc, faset), in_order_of(la, faset)};
  std::vector<std::string> fb_opts = {in_order_of(lc, fbset), in_order_of(lb, fbset)};
  std::vector<std::string> ct_opts = {in_order_of(la, cset), in_order_of(lb, cset)};
  Plan best{};
  best.cost = 1e300;
  const double ab = double(A.bytes()), bb = double(B.bytes()), cb = double(M) * N * A.esz();
  for (auto& fa : fa_opts)
    for (auto& fb : fb_opts)
      for (auto& ct : ct_opts) {
        Plan p{fa, fb, ct, false, false, false, false, false, 0};
        if (is_concat(la, fa, ct)) { p.a_kfast = false; }
        else if (is_concat(la, ct, fa)) { p.a_kfast = true; }
        else { p.permA = true; p.a_kfast = false; p.cost += 2 * ab; }
        if (is_concat(lb, ct, fb)) { p.b_kfast = true; }
        else if (is_concat(lb, fb, ct)) { p.b_kfast = false; }
        else { p.permB = true; p.b_kfast = true; p.cost += 2 * bb; }
        p.permC = (fa + fb) != lc;
        if (p.permC) p.cost += 2 * cb;
        // prefer k-fast/k-fast on ties (coalesced along k for both tiles)
        p.cost += (p.a_kfast ? 0 : 1) + (p.b_kfast ? 0 : 1);
        if (p.cost < best.cost) best = p;
      }

  Tensor Ap = A, Bp = B;
  if (best.permA) Ap = permute(c, A, perm_to(la, best.fa + best.ct));
  if (best.permB) Bp = permute(c, B, perm_to(lb, best.ct + best.fb));
  std::vector<int64_t> cd;
  for (char ch : best.fa + best.fb) cd.push_back(dimof[(unsigned char)ch]);
  Tensor C = c.alloc(cd, A.cplx);
  gemm(c, A.cplx, M, N, K, Ap.data(), best.a_kfast, conjA, Bp.data(), best.b_kfast, conjB,
       C.data(), 0.0);
  if (best.permC) C = permute(c, C, perm_to(best.fa + best.fb, lc));
  std::vector<int64_t> od;
  for (char ch : lc) od.push_back(dimof[(unsigned char)ch]);
  return C.reshaped(od);
}

Tensor matmul(Ctx& c, const Tensor& A, const Tensor& B, bool conjA, bool conjB, bool transA,
              bool transB) {
  PEPS_CHECK(A.rank() == 2 && B.rank() == 2, PS_ESHAPE, "matmul: matrices expected");
  PEPS_CHECK(A.cplx == B.cplx, PS_EARG, "matmul: mixed real/complex");
  int64_t M = transA ? A.dims[1] : A.dims[0];
  int64_t K = transA ? A.dims[0] : A.dims[1];
  int64_t K2 = transB ? B.dims[1] : B.dims[0];
  int64_t N = transB ? B.dims[0] : B.dims[1];
  PEPS_CHECK(K == K2, PS_ESHAPE, "matmul: inner dimension mismatch");
  Tensor C = c.alloc({M, N}, A.cplx);
  gemm(c, A.cplx, M, N, K, A.data(), transA, conjA, B.data(), !transB, conjB, C.data(), 0.0);
  return C;
}

}  // namespace peps
