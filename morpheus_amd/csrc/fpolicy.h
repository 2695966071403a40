// Field policies: the Goldilocks and almost-Goldilocks fields (and their quadratic
// extensions) behind one static interface, so a kernel written once as a template serves
// both crates' symbols. Every operation takes and returns canonical residues, except the
// loads, which canonicalize whatever 64-bit words they read.
#pragma once

#include "field.h"
#include "fieldq.h"

struct FieldGL {
    using E = ext2;
    static MAM_HD u64 canon(u64 a) { return gl_canon(a); }
    static MAM_HD u64 add(u64 a, u64 b) { return gl_add(a, b); }
    static MAM_HD u64 sub(u64 a, u64 b) { return gl_sub(a, b); }
    static MAM_HD u64 mul(u64 a, u64 b) { return gl_mul(a, b); }
    static MAM_HD u64 neg(u64 a) { return gl_neg(a); }
    static MAM_HD u64 sqr(u64 a) { return gl_sqr(a); }
    static MAM_HD u64 dbl(u64 a) { return gl_double(a); }
    static MAM_HD u64 inv(u64 a) { return gl_inv(a); }
    static MAM_HD u64 pow(u64 a, u64 e) { return gl_pow(a, e); }
    static MAM_HD E e(u64 a, u64 b) { return ext2{a, b}; }
    static MAM_HD E e_zero() { return e2_zero(); }
    static MAM_HD E e_one() { return e2_one(); }
    static MAM_HD E e_base(u64 a) { return e2_base(a); }
    static MAM_HD E e_load(const u64 *p, size_t i) { return e2_load(p, i); }
    static MAM_HD void e_store(u64 *p, size_t i, E v) { e2_store(p, i, v); }
    static MAM_HD E e_add(E a, E b) { return e2_add(a, b); }
    static MAM_HD E e_sub(E a, E b) { return e2_sub(a, b); }
    static MAM_HD E e_mul(E a, E b) { return e2_mul(a, b); }
    static MAM_HD E e_neg(E a) { return e2_neg(a); }
    static MAM_HD E e_sqr(E a) { return e2_sqr(a); }
    static MAM_HD E e_inv(E a) { return e2_inv(a); }
    static MAM_HD E e_pow(E a, u64 x) { return e2_pow(a, x); }
    static MAM_HD E e_scale(u64 s, E a) { return e2_scale(s, a); }
    static MAM_HD E e_frob(E a) { return e2_frobenius(a); }
};

struct FieldQ {
    using E = qe2;
    static MAM_HD u64 canon(u64 a) { return fq::canon(a); }
    static MAM_HD u64 add(u64 a, u64 b) { return fq::add(a, b); }
    static MAM_HD u64 sub(u64 a, u64 b) { return fq::sub(a, b); }
    static MAM_HD u64 mul(u64 a, u64 b) { return fq::mul(a, b); }
    static MAM_HD u64 neg(u64 a) { return fq::neg(a); }
    static MAM_HD u64 sqr(u64 a) { return fq::sqr(a); }
    static MAM_HD u64 dbl(u64 a) { return fq::dbl(a); }
    static MAM_HD u64 inv(u64 a) { return fq::inv(a); }
    static MAM_HD u64 pow(u64 a, u64 e) { return fq::pow(a, e); }
    static MAM_HD E e(u64 a, u64 b) { return qe2{a, b}; }
    static MAM_HD E e_zero() { return q2_zero(); }
    static MAM_HD E e_one() { return q2_one(); }
    static MAM_HD E e_base(u64 a) { return q2_base(a); }
    static MAM_HD E e_load(const u64 *p, size_t i) { return q2_load(p, i); }
    static MAM_HD void e_store(u64 *p, size_t i, E v) { q2_store(p, i, v); }
    static MAM_HD E e_add(E a, E b) { return q2_add(a, b); }
    static MAM_HD E e_sub(E a, E b) { return q2_sub(a, b); }
    static MAM_HD E e_mul(E a, E b) { return q2_mul(a, b); }
    static MAM_HD E e_neg(E a) { return q2_neg(a); }
    static MAM_HD E e_sqr(E a) { return q2_sqr(a); }
    static MAM_HD E e_inv(E a) { return q2_inv(a); }
    static MAM_HD E e_pow(E a, u64 x) { return q2_pow(a, x); }
    static MAM_HD E e_scale(u64 s, E a) { return q2_scale(s, a); }
    static MAM_HD E e_frob(E a) { return q2_conj(a); }
};
