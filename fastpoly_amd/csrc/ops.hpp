// Callers of the path built on the device transforms and series:
// poly::pow (power.hpp:17-87) and poly::divmod / div / mod (division.hpp:12-104).
//
// Every result is unique (a^k mod x^n; the Euclidean quotient and remainder),
// so each is computed by one exact route on the device: pow by binary
// powering with truncated products when the reference does (k <= 64, or more
// than p coefficients), else exp(k log(a / c)) c^k; the reference's one- and
// two-term shortcuts give the same limbs and take the same routes here.
// divmod always uses the reversed-inverse quotient (the reference's long
// division for short operands gives the same q and r).  The host reads only
// scan words and output lengths.
#pragma once

namespace fpxg {

static void copy_scaled(Context& c, const HostField& T, uint32_t* dst, const uint32_t* src, size_t n, uint32_t factor) {
  emit_ew(c, T, EW_SCALE, dst, src, nullptr, n, 0, factor);
}

// poly::pow (power.hpp:17-87): out[0, n) = a^k mod x^n.
static void pow_dev(Context& c, const HostField& T, const uint32_t* a, size_t la, uint64_t k, size_t n,
                    uint32_t* out) {
  if (n == 0) return;
  const uint32_t one = T.from_int(1);
  la = std::min(la, n);  // coefficients at or above x^n never reach the result
  if (k == 0 || la == 0) {  // :20-24, and a == 0
    emit_ew(c, T, EW_FILL, out, nullptr, nullptr, n, 0, 0);
    if (k == 0) emit_ew(c, T, EW_FILL, out, nullptr, nullptr, 1, 0, one);
    return;
  }
  if (k == 1) {  // :25-29
    copy_scaled(c, T, out, a, la, one);
    if (la < n) emit_ew(c, T, EW_FILL, out + la, nullptr, nullptr, n - la, 0, 0);
    return;
  }
  const ScanOut s = scan_inputs(c, {ScanRec{a, (uint32_t)la, 0}})[0];
  const size_t v = s.first;
  emit_ew(c, T, EW_FILL, out, nullptr, nullptr, n, 0, 0);
  if (v >= la) return;                         // :30-33
  if (v != 0 && k > (n - 1) / v) return;       // :34
  const size_t sh = (size_t)(v * k), inner = n - sh;
  const size_t lu = std::min(la - v, inner);
  const uint32_t* u = a + v;
  if (k <= 64 || inner > T.p) {  // binary powering with truncated products (:62-73)
    DevBuf rb = c.buf(inner), bb = c.buf(inner), tb = c.buf(inner);
    uint32_t *r = rb.get(), *base = bb.get(), *tmp = tb.get();
    emit_ew(c, T, EW_FILL, r, nullptr, nullptr, 1, 0, one);
    copy_scaled(c, T, base, u, lu, one);
    size_t lr = 1, lbs = lu;
    for (uint64_t e = k; e; e >>= 1) {
      if (e & 1) {
        const size_t m = std::min(lr + lbs - 1, inner);
        conv_dev(c, T, r, lr, base, lbs, tmp, m);
        std::swap(r, tmp);
        lr = m;
      }
      if (e > 1) {
        const size_t m = std::min(2 * lbs - 1, inner);
        conv_dev(c, T, base, lbs, base, lbs, tmp, m);
        std::swap(base, tmp);
        lbs = m;
      }
    }
    copy_scaled(c, T, out + sh, r, lr, one);
    return;
  }
  // exp(k * log(a / c)) * c^k (:75-86); inner <= p here, as log and exp need.
  const uint32_t c0 = s.afirst;
  DevBuf nb = c.buf(lu), lb = c.buf(inner);
  copy_scaled(c, T, nb.get(), u, lu, T.inv(c0));
  series_dev(c, T, FPXG_LOG, nb.get(), lu, inner, lb.get());
  copy_scaled(c, T, lb.get(), lb.get(), inner, T.from_int(k));
  series_dev(c, T, FPXG_EXP, lb.get(), inner, inner, out + sh);
  copy_scaled(c, T, out + sh, out + sh, inner, T.pow(c0, k));
}

// poly::divmod (division.hpp:12-104): a = b q + r with deg r < deg b, both
// trimmed.  q (capacity q_cap) or r (capacity r_cap) may be null.
static void divmod_dev(Context& c, const HostField& T, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
                       uint32_t* q, size_t q_cap, size_t* q_len, uint32_t* r, size_t r_cap, size_t* r_len) {
  const std::vector<ScanOut> s = scan_inputs(c, {ScanRec{a, (uint32_t)la, 0}, ScanRec{b, (uint32_t)lb, 0}});
  const size_t an = s[0].last1, bn = s[1].last1;  // trailing zeros trimmed (:14-16)
  *q_len = *r_len = 0;
  if (bn == 0) throw Error(E_DOMAIN, "poly::divmod: division by zero polynomial");
  const uint32_t one = T.from_int(1);
  if (an < bn) {  // :18-22
    if (r) {
      if (r_cap < an) throw Error(E_ARG, "remainder capacity too small");
      if (an) copy_scaled(c, T, r, a, an, one);
    }
    *r_len = an;
    return;
  }
  const size_t dn = an - bn + 1;
  if (q && q_cap < dn) throw Error(E_ARG, "quotient capacity too small");
  if (bn == 1) {  // :23-28
    if (q) copy_scaled(c, T, q, a, an, T.inv(s[1].a0));
    *q_len = an;
    return;
  }
  // q = rev(rev(a) inv(rev(b)) mod x^dn) (:53-62)
  const size_t rn = bn - 1, lrb = std::min(bn, dn);
  DevBuf ra = c.buf(dn), rbv = c.buf(lrb), ib = c.buf(dn), rq = c.buf(dn), qb = c.buf(dn);
  emit_ew(c, T, EW_REVERSE, ra.get(), a + (an - dn), nullptr, dn, 0, 0);
  emit_ew(c, T, EW_REVERSE, rbv.get(), b + (bn - lrb), nullptr, lrb, 0, 0);
  series_dev(c, T, FPXG_INV, rbv.get(), lrb, dn, ib.get());
  conv_dev(c, T, ra.get(), dn, ib.get(), dn, rq.get(), dn);
  uint32_t* qd = q ? q : qb.get();
  emit_ew(c, T, EW_REVERSE, qd, rq.get(), nullptr, dn, 0, 0);
  *q_len = dn;
  // r = (a - b q) mod x^rn, trimmed (:63-75)
  DevBuf pr = c.buf(rn), rr = c.buf(rn);
  conv_dev(c, T, b, bn, qd, dn, pr.get(), rn);
  emit_ew(c, T, EW_SUB_FROM, rr.get(), a, pr.get(), rn, (uint32_t)std::min(la, rn), 0);
  const size_t lr = scan_inputs(c, {ScanRec{rr.get(), (uint32_t)rn, 0}})[0].last1;
  if (r) {
    if (r_cap < lr) throw Error(E_ARG, "remainder capacity too small");
    if (lr) copy_scaled(c, T, r, rr.get(), lr, one);
  }
  *r_len = lr;
}

}  // namespace fpxg
