// Per-size implementation of the step schedules: every kernel launch of the fused FP32 steps
// (spectral and reference rules, single handles and batched ensemble launches), the FP64 plain
// path, the transforms used by the adjoint, the energy reductions in k-space. Compiled once per
// grid edge with -DGSK_IMPL_N=<n>; the impl_* entry points of that size are instantiated here.
#include "gsk_internal.h"

#ifndef GSK_IMPL_N
#error "compile gsk_impl.hip with -DGSK_IMPL_N=<grid edge>"
#endif

namespace {

// ------------------------------------------------------------------ per-size implementation
template <int N>
struct Impl {
  static constexpr int M = N / 2;

  static int set_attrs() {
    CHK(lds_limit((const void*)k_zfwd<M, Z_COPY>));
    CHK(lds_limit((const void*)k_zfwd<M, Z_SQUARE>));
    CHK(lds_limit((const void*)k_zfwd_local<M>));
    CHK(lds_limit((const void*)k_zfwd_local_r<M, 2>));
    CHK(lds_limit((const void*)k_zfwd_local_r<M, 3>));
    CHK(lds_limit((const void*)k_zfwd_local_r<M, 4>));
    CHK(lds_limit((const void*)k_zfwd_local_r<M, 5>));
    CHK(lds_limit((const void*)k_zmid_mart<M, 0>));
    CHK(lds_limit((const void*)k_zmid_mart<M, 1>));
    CHK(lds_limit((const void*)k_zmid_mart<M, 2>));
    CHK(lds_limit((const void*)k_zmid_mart<M, 3>));
    CHK(lds_limit((const void*)k_zmid_mart<M, 4>));
    CHK(lds_limit((const void*)k_zinv<M>));
    CHK(lds_limit((const void*)k_zinv_sq<M>));
    CHK(lds_limit((const void*)k_zinv_local_r<M, 2>));
    CHK(lds_limit((const void*)k_zinv_local_r<M, 3>));
    CHK(lds_limit((const void*)k_zinv_local_r<M, 4>));
    CHK(lds_limit((const void*)k_zinv_local_r<M, 5>));
    CHK(lds_limit((const void*)k_ypass<N, true>));
    CHK(lds_limit((const void*)k_ypass<N, false>));
    CHK(lds_limit((const void*)k_xpass<N, true>));
    CHK(lds_limit((const void*)k_xpass<N, false>));
    CHK(lds_limit((const void*)k_xcontract<N>));
    CHK(lds_limit((const void*)k_xcontract_r<N, 1>));
    CHK(lds_limit((const void*)k_xcontract_r<N, 2>));
    CHK(lds_limit((const void*)k_xcontract_r<N, 3>));
    CHK(lds_limit((const void*)k_xcontract_r<N, 4>));
    CHK(lds_limit((const void*)k_xupdate_mart<N, 1, 0>));
    CHK(lds_limit((const void*)k_xupdate_mart<N, 2, 0>));
    CHK(lds_limit((const void*)k_xupdate_mart<N, 1, 1>));
    CHK(lds_limit((const void*)k_xupdate_mart<N, 2, 1>));
    CHK(lds_limit((const void*)k_xupdate_mart<N, 3, 0>));
    CHK(lds_limit((const void*)k_xupdate_mart<N, 3, 1>));
    CHK(lds_limit((const void*)k_xupdate_prec<N, 1, 0>));
    CHK(lds_limit((const void*)k_xupdate_prec<N, 2, 0>));
    CHK(lds_limit((const void*)k_xupdate_prec<N, 1, 1>));
    CHK(lds_limit((const void*)k_xupdate_prec<N, 2, 1>));
    CHK(lds_limit((const void*)r_xenergy_mart<N>));
    CHK(lds_limit((const void*)r_xenergy_mart_r<N, 1>));
    CHK(lds_limit((const void*)r_xenergy_mart_r<N, 2>));
    CHK(lds_limit((const void*)r_xenergy_mart_r<N, 3>));
    CHK(lds_limit((const void*)r_xenergy_mart_r<N, 4>));
    CHK(lds_limit((const void*)p_zrows<double2, N, true>));
    CHK(lds_limit((const void*)p_zrows<double2, N, false>));
    CHK(lds_limit((const void*)p_strided<double2, N, true>));
    CHK(lds_limit((const void*)p_strided<double2, N, false>));
    CHK(lds_limit((const void*)p_zrows<float2, N, true>));
    CHK(lds_limit((const void*)p_zrows<float2, N, false>));
    CHK(lds_limit((const void*)p_strided<float2, N, true>));
    CHK(lds_limit((const void*)p_strided<float2, N, false>));
    // the batched ensemble instantiations (member table)
    CHK(lds_limit((const void*)k_ypass<N, true, true>));
    CHK(lds_limit((const void*)k_ypass<N, false, true>));
    CHK(lds_limit((const void*)k_xcontract_r<N, 1, true>));
    CHK(lds_limit((const void*)k_xcontract_r<N, 2, true>));
    CHK(lds_limit((const void*)k_xcontract_r<N, 3, true>));
    CHK(lds_limit((const void*)k_xcontract_r<N, 4, true>));
    CHK(lds_limit((const void*)k_xupdate_mart<N, 1, 0, true>));
    CHK(lds_limit((const void*)k_zmid_mart<M, 1, true>));
    CHK(lds_limit((const void*)k_zmid_mart<M, 2, true>));
    CHK(lds_limit((const void*)k_zmid_mart<M, 3, true>));
    CHK(lds_limit((const void*)k_zmid_mart<M, 4, true>));
    CHK(lds_limit((const void*)k_zinv_sq<M, true>));
    if constexpr (split_size<N>()) {
      CHK(lds_limit((const void*)k_xcontract_split<N, 1, 0>));
      CHK(lds_limit((const void*)k_xcontract_split<N, 2, 0>));
      CHK(lds_limit((const void*)k_xcontract_split<N, 3, 0>));
      CHK(lds_limit((const void*)k_xcontract_split<N, 4, 0>));
      CHK(lds_limit((const void*)k_xcontract_split<N, 1, 1>));
      CHK(lds_limit((const void*)k_xcontract_split<N, 2, 1>));
      CHK(lds_limit((const void*)k_xcontract_split<N, 3, 1>));
      CHK(lds_limit((const void*)k_xcontract_split<N, 4, 1>));
      CHK(lds_limit((const void*)k_xupdate_split<N, 0>));
      CHK(lds_limit((const void*)k_xupdate_split<N, 1>));
      CHK(lds_limit((const void*)k_sub_mid<N, 1>));
      CHK(lds_limit((const void*)k_sub_mid<N, 2>));
      CHK(lds_limit((const void*)k_sub_mid<N, 3>));
      CHK(lds_limit((const void*)k_sub_mid<N, 4>));
      CHK(lds_limit((const void*)k_sub_sq<N>));
      CHK(lds_limit((const void*)k_sub_sq<N, false>));
      CHK(lds_limit((const void*)k_sub_fwd<N, 0>));
      CHK(lds_limit((const void*)k_sub_fwd<N, 1>));
      CHK(lds_limit((const void*)k_xupdate_split_prec<N, 0>));
      CHK(lds_limit((const void*)k_xupdate_split_prec<N, 1>));
      CHK(lds_limit((const void*)k_sub_local<N, 2>));
      CHK(lds_limit((const void*)k_sub_local<N, 3>));
      CHK(lds_limit((const void*)k_sub_local<N, 4>));
      CHK(lds_limit((const void*)k_sub_local<N, 5>));
    }
    if constexpr (plane_size<N>()) {
      CHK(lds_limit((const void*)k_plane_sq<N, true>));
      CHK(lds_limit((const void*)k_plane_mid<N, 1, true>));
      CHK(lds_limit((const void*)k_plane_mid<N, 2, true>));
      CHK(lds_limit((const void*)k_plane_mid<N, 3, true>));
      CHK(lds_limit((const void*)k_plane_mid<N, 4, true>));
      CHK(lds_limit((const void*)k_plane_sq<N>));
      CHK(lds_limit((const void*)k_plane_sq0<N>));
      CHK(lds_limit((const void*)k_plane_mid<N, 1>));
      CHK(lds_limit((const void*)k_plane_mid<N, 2>));
      CHK(lds_limit((const void*)k_plane_mid<N, 3>));
      CHK(lds_limit((const void*)k_plane_mid<N, 4>));
    }
    return GSK_OK;
  }

  // y inverse of eta, last z pass (eta written, squared) and y forward in one launch, one x-plane
  // of one field per work group (plane_on); adv: the step counter the launch advances
  static int plane_sq(gsk_sim* s, uint32_t* adv) {
    if constexpr (plane_size<N>()) {
      LAUNCH((k_plane_sq<N>), dim3(N, s->F), plane_threads<N>(), plane_lds<N>(), s, s->kbuf, s->fld, geo_next(s), s->twN,
             s->twM, s->rtn, adv, (const MemberPtrs*)nullptr);
      return GSK_OK;
    } else {
      (void)s; (void)adv;
      return fail(GSK_ERR_UNSUPPORTED, "the plane pass takes n = 32, 64, 128");
    }
  }

  // y inverse of R, middle z pass and y forward of F in one launch, one x-plane of one variant
  // per work group (plane_on)
  static int plane_mid(gsk_sim* s, const MartK& mk) {
    if constexpr (plane_size<N>()) {
      const dim3 grid(N, s->F);
      const Geo g = geo_next(s);
      constexpr int T = plane_threads<N>();
      constexpr size_t lds = plane_lds<N>();
      switch (s->F) {
        case 1: LAUNCH((k_plane_mid<N, 1>), grid, T, lds, s, s->kbuf, s->fld, g, mk, s->twN, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
        case 2: LAUNCH((k_plane_mid<N, 2>), grid, T, lds, s, s->kbuf, s->fld, g, mk, s->twN, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
        case 3: LAUNCH((k_plane_mid<N, 3>), grid, T, lds, s, s->kbuf, s->fld, g, mk, s->twN, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
        case 4: LAUNCH((k_plane_mid<N, 4>), grid, T, lds, s, s->kbuf, s->fld, g, mk, s->twN, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
        default: return fail(GSK_ERR_UNSUPPORTED, "the plane pass takes 1..4 variants");
      }
      return GSK_OK;
    } else {
      (void)s; (void)mk;
      return fail(GSK_ERR_UNSUPPORTED, "the plane pass takes n = 32, 64, 128");
    }
  }

  // ---- fused building blocks
  // LDS of a z kernel holding nf field tiles
  static size_t zlds(const gsk_sim* s, int nf) { return (2 * M + (size_t)nf * s->g.L * zpitch(M)) * sizeof(float2); }
  // shape of the middle martensite pass (its own rows per tile)
  static constexpr int LZ = zmid_rows(M), ZMT = zmid_threads<M>();
  static constexpr size_t ZM_LDS = (2 * M + (size_t)LZ * zpitch_l(M, LZ)) * sizeof(float2);
  static int zfwd(gsk_sim* s, int op, const float* in, int F) {
    const int rows = s->n * s->n / s->g.L;
    if (op == Z_COPY) {
      LAUNCH((k_zfwd<M, Z_COPY>), dim3(rows, F), s->zthreads, zlds(s, 1), s, in, s->kbuf, geo_next(s), s->twM, s->rtn);
    } else if (op == Z_SQUARE) {
      LAUNCH((k_zfwd<M, Z_SQUARE>), dim3(rows, F), s->zthreads, zlds(s, 1), s, in, s->kbuf, geo_next(s), s->twM, s->rtn);
    } else {
      PrecK pk = prec_consts(s);
      const bool ref = s->cfg.rule == GSK_RULE_REFERENCE;
      const float* R = ref ? (const float*)s->rbuf : (const float*)nullptr;
      const float el = ref ? (float)(2.0 * s->cfg.xi) : 0.f, cn = ref ? (float)s->cfg.noise : 0.f;
      switch (F) {  // every field of a tile in registers for up to 5 fields
        case 2: LAUNCH((k_zfwd_local_r<M, 2>), rows, s->zthreads, zlds(s, 1), s, in, s->kbuf, s->g, pk, s->twM, s->rtn, R, el, cn); break;
        case 3: LAUNCH((k_zfwd_local_r<M, 3>), rows, s->zthreads, zlds(s, 1), s, in, s->kbuf, s->g, pk, s->twM, s->rtn, R, el, cn); break;
        case 4: LAUNCH((k_zfwd_local_r<M, 4>), rows, s->zthreads, zlds(s, 1), s, in, s->kbuf, s->g, pk, s->twM, s->rtn, R, el, cn); break;
        case 5: LAUNCH((k_zfwd_local_r<M, 5>), rows, s->zthreads, zlds(s, 1), s, in, s->kbuf, s->g, pk, s->twM, s->rtn, R, el, cn); break;
        default: LAUNCH((k_zfwd_local<M>), rows, s->zthreads, zlds(s, F), s, in, s->kbuf, s->g, pk, s->twM, s->rtn, R, el, cn);
      }
    }
    return GSK_OK;
  }
  static int ypass(gsk_sim* s, bool fwd, int F) {
    const dim3 grid(s->g.ntile, GSK_Y_PERFIELD ? F : 1);
    const Geo g = GSK_Y_PERFIELD ? geo_next(s) : s->g;
    if (fwd) LAUNCH((k_ypass<N, true>), grid, s->xthreads, s->lds_y, s, s->kbuf, g, F, s->twN, (const MemberPtrs*)nullptr);
    else LAUNCH((k_ypass<N, false>), grid, s->xthreads, s->lds_y, s, s->kbuf, g, F, s->twN, (const MemberPtrs*)nullptr);
    return GSK_OK;
  }
  static int zinv(gsk_sim* s, float* out, int F, uint32_t* adv = nullptr) {
    const int rows = s->n * s->n / s->g.L;
    LAUNCH((k_zinv<M>), dim3(rows, F), s->zthreads, zlds(s, 1), s, s->kbuf, out, geo_next(s), s->twM, s->rtn, adv);
    return GSK_OK;
  }

  // k-space updates in groups of 2 fields (register-resident x transforms)
  template <int MODE>
  static int xupdate_mart(gsk_sim* s, const MartK& mk) {
    const int gx = s->g.ntile;
    if (GSK_XUPD_G == 0) {  // one field per work group, one launch
      LAUNCH((k_xupdate_mart<N, 1, MODE>), dim3(gx, s->F), s->xthreads, s->lds_y, s, s->kbuf, s->state, geo_next(s), mk,
             s->twN, 0, (const MemberPtrs*)nullptr);
      return GSK_OK;
    }
    for (int f0 = 0; f0 < s->F;) {
      const int G = xupd_group(s->F - f0);
      if (G == 3)
        LAUNCH((k_xupdate_mart<N, 3, MODE>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->state, s->g, mk, s->twN, f0, (const MemberPtrs*)nullptr);
      else if (G == 2)
        LAUNCH((k_xupdate_mart<N, 2, MODE>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->state, s->g, mk, s->twN, f0, (const MemberPtrs*)nullptr);
      else
        LAUNCH((k_xupdate_mart<N, 1, MODE>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->state, s->g, mk, s->twN, f0, (const MemberPtrs*)nullptr);
      f0 += G;
    }
    return GSK_OK;
  }
  template <int MODE>
  static int xupdate_prec(gsk_sim* s, const PrecK& pk) {
    const int gx = s->g.ntile;
    if (GSK_XUPD_G == 0) {  // one field per work group, one launch
      LAUNCH((k_xupdate_prec<N, 1, MODE>), dim3(gx, s->F), s->xthreads, s->lds_y, s, s->kbuf, s->state, s->Bt,
             geo_next(s), pk, s->twN, 0);
      return GSK_OK;
    }
    PrecK pg = pk;
    for (int f0 = 0; f0 < s->F; f0 += 2) {
      if (s->F - f0 >= 2)
        LAUNCH((k_xupdate_prec<N, 2, MODE>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->state, s->Bt, s->g, pg, s->twN,
               f0);
      else
        LAUNCH((k_xupdate_prec<N, 1, MODE>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->state, s->Bt, s->g, pg, s->twN,
               f0);
      pg.adv = nullptr;  // one advance per step
    }
    return GSK_OK;
  }

  // ---- split-y schedule (n = 256): sub-plane launches have grid (n * P sub-planes, fields)
  static constexpr size_t SPLIT_XLDS = (size_t)(N + N * split_w<kSplitWc>()) * sizeof(float2);   // x update
  static constexpr size_t SPLIT_CLDS = (size_t)(N + N * split_w<kSplitWcC>()) * sizeof(float2);  // contraction
  // E-form spectra of `count` real fields at `in` (square: of their squares) into the k-space
  // buffer's first `count` fields; the handle's fields by default
  static int split_fwd(gsk_sim* s, bool square, const float* in = nullptr, int count = 0) {
    if constexpr (split_size<N>()) {
      if (!in) in = s->fld;
      const dim3 grid(N * kSplitP, count > 0 ? count : s->F);
      if (square)
        LAUNCH((k_sub_fwd<N, 1>), grid, sub_threads<N>(), sub_lds<N>(), s, in, s->kbuf, s->sg, s->twN, s->twM, s->rtn);
      else
        LAUNCH((k_sub_fwd<N, 0>), grid, sub_threads<N>(), sub_lds<N>(), s, in, s->kbuf, s->sg, s->twN, s->twM, s->rtn);
      return GSK_OK;
    } else {
      (void)s; (void)square; (void)in; (void)count;
      return fail(GSK_ERR_UNSUPPORTED, "the split-y schedule takes n = 128, 192 or 256");
    }
  }
  // the contraction over the first `count` spectra of the k-space buffer (the handle's F by default)
  template <int MODE>
  static int split_xcontract(gsk_sim* s, const MartK& mk, int count = 0) {
    if constexpr (split_size<N>()) {
      const int gx = s->sg.ntile_c;
      if (MODE == 1 && gx > s->rblocks) return fail(GSK_ERR, "reduction scratch too small");
      constexpr int T = split_cthreads<N>();
      switch (count > 0 ? count : s->F) {
        case 1: LAUNCH((k_xcontract_split<N, 1, MODE>), gx, T, SPLIT_CLDS, s, s->kbuf, s->Bt, s->sg, mk, s->twN, s->part); break;
        case 2: LAUNCH((k_xcontract_split<N, 2, MODE>), gx, T, SPLIT_CLDS, s, s->kbuf, s->Bt, s->sg, mk, s->twN, s->part); break;
        case 3: LAUNCH((k_xcontract_split<N, 3, MODE>), gx, T, SPLIT_CLDS, s, s->kbuf, s->Bt, s->sg, mk, s->twN, s->part); break;
        default: LAUNCH((k_xcontract_split<N, 4, MODE>), gx, T, SPLIT_CLDS, s, s->kbuf, s->Bt, s->sg, mk, s->twN, s->part);
      }
      return GSK_OK;
    } else {
      (void)s; (void)mk; (void)count;
      return fail(GSK_ERR_UNSUPPORTED, "the split-y schedule takes n = 128, 192 or 256");
    }
  }
  template <int MODE>
  static int split_xupdate(gsk_sim* s, const MartK& mk) {
    if constexpr (split_size<N>()) {
      LAUNCH((k_xupdate_split<N, MODE>), dim3(s->sg.ntile, s->F), split_xthreads<N>(), SPLIT_XLDS, s, s->kbuf,
             s->state, s->sg, mk, s->twN);
      return GSK_OK;
    } else {
      (void)s; (void)mk;
      return fail(GSK_ERR_UNSUPPORTED, "the split-y schedule takes n = 128, 192 or 256");
    }
  }
  static int step_mart_split(gsk_sim* s) {
    if constexpr (split_size<N>()) {
      const MartK mk = mart_consts(s);
      // resident sub-plane work groups: two per CU, k_sub_mid's count a multiple of 8 F
      static_assert((N * kSplitP) % 8 == 0, "sub-planes in groups of 8");
      if (!s->cus) HIPCHK(hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, s->dev));
      const int nitems = N * kSplitP * s->F, g8 = 8 * s->F;
      const int gm = std::min(nitems, std::max(g8, sub_groups_per_cu<N>() * s->cus / g8 * g8));
      const int gq = std::min(nitems, std::max(8, sub_groups_per_cu<N>() * s->cus / 8 * 8));
      if (!s->sq_ready) CHK(split_fwd(s, true));
      CHK(split_xcontract<0>(s, mk));
      switch (s->F) {
        case 1: LAUNCH((k_sub_mid<N, 1>), gm, sub_threads<N>(), sub_lds<N>(), s, s->kbuf, s->fld, s->sg, mk, s->twN, s->twM, s->rtn); break;
        case 2: LAUNCH((k_sub_mid<N, 2>), gm, sub_threads<N>(), sub_lds<N>(), s, s->kbuf, s->fld, s->sg, mk, s->twN, s->twM, s->rtn); break;
        case 3: LAUNCH((k_sub_mid<N, 3>), gm, sub_threads<N>(), sub_lds<N>(), s, s->kbuf, s->fld, s->sg, mk, s->twN, s->twM, s->rtn); break;
        default: LAUNCH((k_sub_mid<N, 4>), gm, sub_threads<N>(), sub_lds<N>(), s, s->kbuf, s->fld, s->sg, mk, s->twN, s->twM, s->rtn);
      }
      CHK(split_xupdate<0>(s, mk));
      LAUNCH((k_sub_sq<N>), gq, sub_threads<N>(), sub_lds<N>(), s, s->kbuf, s->fld, s->sg, s->twN, s->twM, s->rtn, s->dstep,
             nitems);
      s->sq_ready = true;
      return GSK_OK;
    } else {
      (void)s;
      return fail(GSK_ERR_UNSUPPORTED, "the split-y schedule takes n = 128, 192 or 256");
    }
  }

  // precipitate, split-y schedule: the update of the F fields (MODE 1: the state from their E-form
  // transforms), the local forces of the F fields into E-form spectra, the inverse into real fields
  template <int MODE>
  static int split_xupdate_prec(gsk_sim* s, const PrecK& pk) {
    if constexpr (split_size<N>()) {
      LAUNCH((k_xupdate_split_prec<N, MODE>), dim3(s->sg.ntile, s->F), split_xthreads<N>(), SPLIT_XLDS, s, s->kbuf,
             s->state, s->Bt, s->sg, pk, s->twN);
      return GSK_OK;
    } else {
      (void)s; (void)pk;
      return fail(GSK_ERR_UNSUPPORTED, "the split-y schedule takes n = 128, 192 or 256");
    }
  }
  static int split_local(gsk_sim* s, const PrecK& pk, const float* R, float el, float cn) {
    if constexpr (split_size<N>()) {
      const int items = N * kSplitP * s->F;
      const SGeo& g = s->sg;
      switch (s->F) {
        case 2: LAUNCH((k_sub_local<N, 2>), items, sub_threads<N>(), sub_lds<N>(), s, s->fld, s->kbuf, g, pk, R, el, cn, s->twN, s->twM, s->rtn); break;
        case 3: LAUNCH((k_sub_local<N, 3>), items, sub_threads<N>(), sub_lds<N>(), s, s->fld, s->kbuf, g, pk, R, el, cn, s->twN, s->twM, s->rtn); break;
        case 4: LAUNCH((k_sub_local<N, 4>), items, sub_threads<N>(), sub_lds<N>(), s, s->fld, s->kbuf, g, pk, R, el, cn, s->twN, s->twM, s->rtn); break;
        default: LAUNCH((k_sub_local<N, 5>), items, sub_threads<N>(), sub_lds<N>(), s, s->fld, s->kbuf, g, pk, R, el, cn, s->twN, s->twM, s->rtn);
      }
      return GSK_OK;
    } else {
      (void)s; (void)pk; (void)R; (void)el; (void)cn;
      return fail(GSK_ERR_UNSUPPORTED, "the split-y schedule takes n = 128, 192 or 256");
    }
  }
  // y inverse and C2R of the first `count` A-form spectra into real fields at out; adv (may be
  // null): the step counter the launch advances
  static int split_inv(gsk_sim* s, float* out, int count, uint32_t* adv) {
    if constexpr (split_size<N>()) {
      if (!s->cus) HIPCHK(hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, s->dev));
      const int nitems = N * kSplitP * count, gq = std::min(nitems, std::max(8, sub_groups_per_cu<N>() * s->cus / 8 * 8));
      LAUNCH((k_sub_sq<N, false>), gq, sub_threads<N>(), sub_lds<N>(), s, s->kbuf, out, s->sg, s->twN, s->twM, s->rtn, adv,
             nitems);
      return GSK_OK;
    } else {
      (void)s; (void)out; (void)count; (void)adv;
      return fail(GSK_ERR_UNSUPPORTED, "the split-y schedule takes n = 128, 192 or 256");
    }
  }
  // Spectral precipitate step, split-y schedule (three launches): x update of the carried local
  // forces (advancing the step counter), y inverse and C2R of the fields, the next step's local
  // forces from the new fields.
  static int step_prec_split(gsk_sim* s) {
    PrecK pk = prec_consts(s);
    if (!s->sq_ready) CHK(split_local(s, pk, nullptr, 0.f, 0.f));
    pk.adv = s->dstep;
    CHK(split_xupdate_prec<0>(s, pk));
    pk.adv = nullptr;
    CHK(split_inv(s, s->fld, s->F, nullptr));
    CHK(split_local(s, pk, nullptr, 0.f, 0.f));
    s->sq_ready = true;
    return GSK_OK;
  }

  static int init_state(gsk_sim* s) {
    if (s->split && s->cfg.physics == GSK_PHYS_PREC) {  // E-form transforms of the fields, then the state
      CHK(split_fwd(s, false));
      PrecK pk = prec_consts(s);
      return split_xupdate_prec<1>(s, pk);
    }
    if (s->split) {  // E-form transform of eta, then the state in the split tile order
      CHK(split_fwd(s, false));
      MartK mk = mart_consts(s);
      return split_xupdate<1>(s, mk);
    }
    if (s->cfg.physics == GSK_PHYS_MART) {
      CHK(zfwd(s, Z_COPY, s->fld, s->F));
      CHK(ypass(s, true, s->F));
      MartK mk = mart_consts(s);
      CHK(xupdate_mart<1>(s, mk));
    } else {
      CHK(zfwd(s, Z_COPY, s->fld, s->F));
      CHK(ypass(s, true, s->F));
      PrecK pk = prec_consts(s);
      CHK(xupdate_prec<1>(s, pk));
    }
    return GSK_OK;
  }

  // z-transformed eta^2 of the current fields into the k-space buffer (y,z-transformed where
  // plane_on): the first kernel of a step, unless the previous step's fused last pass has left
  // it there
  static int square_fwd(gsk_sim* s) {
    if (s->split && s->cfg.physics == GSK_PHYS_PREC) {
      CHK(split_local(s, prec_consts(s), nullptr, 0.f, 0.f));
    } else if (s->split) {
      CHK(split_fwd(s, true));
    } else if (plane_on(s)) {
      if constexpr (plane_size<N>()) {
        LAUNCH((k_plane_sq0<N>), dim3(N, s->F), plane_threads<N>(), plane_lds<N>(), s, s->fld, s->kbuf, geo_next(s),
               s->twN, s->twM, s->rtn);
      }
    } else {
      CHK(zfwd(s, s->cfg.physics == GSK_PHYS_MART ? Z_SQUARE : Z_PREC_LOCAL, s->fld, s->F));
    }
    s->sq_ready = true;
    return GSK_OK;
  }

  static int step_mart_fused(gsk_sim* s) {
    if (s->split) return step_mart_split(s);
    const int gx = s->g.ntile;
    const int F = s->F;
    MartK mk = mart_consts(s);
    const bool pl = plane_on(s);
    if (!s->sq_ready) CHK(square_fwd(s));
    if (!pl) CHK(ypass(s, true, F));
    switch (F) {  // register-resident contraction for up to 4 variants
      case 1: LAUNCH((k_xcontract_r<N, 1>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      case 2: LAUNCH((k_xcontract_r<N, 2>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      case 3: LAUNCH((k_xcontract_r<N, 3>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      case 4: LAUNCH((k_xcontract_r<N, 4>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      default: LAUNCH((k_xcontract<N>), gx, s->xthreads, s->lds_xc, s, s->kbuf, s->Bt, s->g, mk, s->twN);
    }
    if (pl) {
      CHK(plane_mid(s, mk));
    } else {
      CHK(ypass(s, false, F));
      {
        const int grid = s->n * s->n / LZ;
        const size_t lds = ZM_LDS;
        switch (F) {  // eta prefetched into registers for up to 4 variants
          case 1: LAUNCH((k_zmid_mart<M, 1>), grid, ZMT, lds, s, s->kbuf, s->fld, s->g, mk, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
          case 2: LAUNCH((k_zmid_mart<M, 2>), grid, ZMT, lds, s, s->kbuf, s->fld, s->g, mk, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
          case 3: LAUNCH((k_zmid_mart<M, 3>), grid, ZMT, lds, s, s->kbuf, s->fld, s->g, mk, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
          case 4: LAUNCH((k_zmid_mart<M, 4>), grid, ZMT, lds, s, s->kbuf, s->fld, s->g, mk, s->twM, s->rtn, (const MemberPtrs*)nullptr); break;
          default: LAUNCH((k_zmid_mart<M, 0>), grid, ZMT, lds, s, s->kbuf, s->fld, s->g, mk, s->twM, s->rtn, (const MemberPtrs*)nullptr);
        }
      }
      CHK(ypass(s, true, F));
    }
    CHK(xupdate_mart<0>(s, mk));
    if (pl) {  // eta back to real space, fused with the next step's forward passes of eta^2
      CHK(plane_sq(s, s->dstep));
    } else {  // eta back to real space, fused with the next step's forward z pass of eta^2
      CHK(ypass(s, false, F));
      const int rows = s->n * s->n / s->g.L;
      LAUNCH((k_zinv_sq<M>), dim3(rows, F), s->zthreads, zlds(s, 1), s, s->kbuf, s->fld, geo_next(s), s->twM, s->rtn,
             s->dstep, (const MemberPtrs*)nullptr);
    }
    s->sq_ready = true;
    return GSK_OK;
  }

  // One spectral martensite step of `cnt` members of this grid size in every launch
  // (grid.z = member). The members' buffers and constants come from the device table `tab`;
  // s0, one of them, supplies the geometry, launch shapes and twiddle tables, which members
  // of one size share. Every member must hold its transformed eta^2 (sq_ready; y,z-transformed
  // where plane_on, which holds for all members of one size alike).
  static int step_mart_batched(gsk_sim* s0, int cnt, const MemberPtrs* tab, hipStream_t st) {
    const int F = s0->F, gx = s0->g.ntile, rows = s0->n * s0->n / s0->g.L;
    const MartK mk{};  // replaced by the table entry in every work group
    const size_t lz = zlds(s0, 1);
    float2* nk = nullptr;
    auto y = [&](bool fwd) -> int {
      if (fwd)
        hipLaunchKernelGGL((k_ypass<N, true, true>), dim3(gx, F, cnt), dim3(s0->xthreads), s0->lds_y, st, nk, geo_next(s0),
                           F, s0->twN, tab);
      else
        hipLaunchKernelGGL((k_ypass<N, false, true>), dim3(gx, F, cnt), dim3(s0->xthreads), s0->lds_y, st, nk, geo_next(s0),
                           F, s0->twN, tab);
      HIPCHK(hipGetLastError());
      return GSK_OK;
    };
    const bool pl = plane_on(s0);
    if (!pl) CHK(y(true));
    switch (F) {
#define GSK_XC(V)                                                                                                  \
  case V:                                                                                                          \
    hipLaunchKernelGGL((k_xcontract_r<N, V, true>), dim3(gx, 1, cnt), dim3(s0->xthreads), s0->lds_y, st, nk,            \
                       (const float*)nullptr, s0->g, mk, s0->twN, tab);                                            \
    break;
      GSK_XC(1) GSK_XC(2) GSK_XC(3) GSK_XC(4)
#undef GSK_XC
      default: return fail(GSK_ERR_UNSUPPORTED, "batched ensemble launches take 1..4 variants");
    }
    HIPCHK(hipGetLastError());
    if (pl) {  // the members' own step uses the plane passes: keep them bit-identical
      if constexpr (plane_size<N>()) {
        switch (F) {
#define GSK_PM(V)                                                                                                  \
  case V:                                                                                                          \
    hipLaunchKernelGGL((k_plane_mid<N, V, true>), dim3(N, F, cnt), dim3(plane_threads<N>()), plane_lds<N>(), st, nk,    \
                       (const float*)nullptr, geo_next(s0), mk, s0->twN, s0->twM, s0->rtn, tab);                   \
    break;
          GSK_PM(1) GSK_PM(2) GSK_PM(3) GSK_PM(4)
#undef GSK_PM
          default: return fail(GSK_ERR_UNSUPPORTED, "batched ensemble launches take 1..4 variants");
        }
        HIPCHK(hipGetLastError());
      }
    } else {
      CHK(y(false));
      switch (F) {
#define GSK_ZM(V)                                                                                                  \
  case V:                                                                                                          \
    hipLaunchKernelGGL((k_zmid_mart<M, V, true>), dim3(s0->n * s0->n / LZ, 1, cnt), dim3(ZMT), ZM_LDS, st, nk,         \
                       (const float*)nullptr, s0->g, mk, s0->twM, s0->rtn, tab);                                   \
    break;
        GSK_ZM(1) GSK_ZM(2) GSK_ZM(3) GSK_ZM(4)
#undef GSK_ZM
        default: return fail(GSK_ERR_UNSUPPORTED, "batched ensemble launches take 1..4 variants");
      }
      HIPCHK(hipGetLastError());
      CHK(y(true));
    }
    hipLaunchKernelGGL((k_xupdate_mart<N, 1, 0, true>), dim3(gx, F, cnt), dim3(s0->xthreads), s0->lds_y, st, nk,
                       (float2*)nullptr, geo_next(s0), mk, s0->twN, 0, tab);
    HIPCHK(hipGetLastError());
    if (pl) {
      if constexpr (plane_size<N>()) {
        hipLaunchKernelGGL((k_plane_sq<N, true>), dim3(N, F, cnt), dim3(plane_threads<N>()), plane_lds<N>(), st, nk,
                           (float*)nullptr, geo_next(s0), s0->twN, s0->twM, s0->rtn, (uint32_t*)nullptr, tab);
      }
    } else {
      CHK(y(false));
      hipLaunchKernelGGL((k_zinv_sq<M, true>), dim3(rows, F, cnt), dim3(s0->zthreads), lz, st, nk, (float*)nullptr,
                         geo_next(s0), s0->twM, s0->rtn, (uint32_t*)nullptr, tab);
    }
    HIPCHK(hipGetLastError());
    return GSK_OK;
  }

  // ---- 3D transforms of F real fields <-> natural-layout half spectra in an explicit buffer
  static int r2c_buf(gsk_sim* s, const float* in, float2* kb, int F) {
    const int rows = s->n * s->n / s->g.L;
    LAUNCH((k_zfwd<M, Z_COPY>), dim3(rows, F), s->zthreads, zlds(s, 1), s, in, kb, s->g, s->twM, s->rtn);
    LAUNCH((k_ypass<N, true>), dim3(s->g.ntile, F), s->xthreads, s->lds_y, s, kb, s->g, F, s->twN, (const MemberPtrs*)nullptr);
    LAUNCH((k_xpass<N, true>), s->g.ntile, s->xthreads, s->lds_y, s, kb, s->g, F, 1.0f, s->twN);
    return GSK_OK;
  }
  static int c2r_buf(gsk_sim* s, float2* kb, float* out, int F) {
    LAUNCH((k_xpass<N, false>), s->g.ntile, s->xthreads, s->lds_y, s, kb, s->g, F, 1.0f, s->twN);
    LAUNCH((k_ypass<N, false>), dim3(s->g.ntile, F), s->xthreads, s->lds_y, s, kb, s->g, F, s->twN, (const MemberPtrs*)nullptr);
    const int rows = s->n * s->n / s->g.L;
    LAUNCH((k_zinv<M>), dim3(rows, F), s->zthreads, zlds(s, 1), s, kb, out, s->g, s->twM, s->rtn,
           (uint32_t*)nullptr);
    return GSK_OK;
  }

  // One reverse step of the martensite update at eta (step index k.step), eta1 = next state.
  // lam (in: dLoss/deta1, out: dLoss/deta); partial sums into part (nb X, nb*V*V G, nb <mu,F>).
  static int adj_step(gsk_sim* s, const float* eta, const float* eta1, float* lam, float* mu, float* rq, float2* kb,
                      float2* k2, const float* bh, double* part, int nb, const AdjK& k) {
    const int V = s->V;
    const long N3 = (long)s->n * s->n * s->n;
    CHK(r2c_buf(s, lam, kb, V));
    CHK(r2c_buf(s, eta1, k2, V));
    LAUNCH(k_adj_dmul<>, nb, 256, 0, s, kb, k2, s->g, k, part);
    CHK(c2r_buf(s, kb, mu, V));
    LAUNCH(k_adj_prep<>, nb, 256, 0, s, eta, mu, rq, V, N3);
    CHK(r2c_buf(s, rq, kb, 2 * V));
    double* pg = part + nb;
    switch (V) {
      case 1: LAUNCH((k_adj_contract<1>), nb, 256, 0, s, kb, bh, s->g, k.inv_n3, pg); break;
      case 2: LAUNCH((k_adj_contract<2>), nb, 256, 0, s, kb, bh, s->g, k.inv_n3, pg); break;
      case 3: LAUNCH((k_adj_contract<3>), nb, 256, 0, s, kb, bh, s->g, k.inv_n3, pg); break;
      case 4: LAUNCH((k_adj_contract<4>), nb, 256, 0, s, kb, bh, s->g, k.inv_n3, pg); break;
      default: return fail(GSK_ERR_UNSUPPORTED, "the adjoint supports up to 4 variants");
    }
    CHK(c2r_buf(s, kb, rq, 2 * V));
    LAUNCH(k_adj_lambda<>, nb, 256, 0, s, eta, mu, rq, lam, N3, k, part + nb + nb * V * V);
    return GSK_OK;
  }

  // R_p = IFFT[sum_q B_pq FFT(eta_q^2)]/N of the V fields at eta into out (real, V fields)
  static int elastic_R_real(gsk_sim* s, const float* eta, float* out) {
    const int V = s->V, gx = s->g.ntile;
    MartK mk = mart_consts(s);
    if (s->split) {  // squared, R2C and y forward -> contraction -> y inverse and C2R (three launches)
      CHK(split_fwd(s, true, eta, V));
      CHK(split_xcontract<0>(s, mk, V));
      return split_inv(s, out, V, nullptr);
    }
    CHK(zfwd(s, Z_SQUARE, eta, V));
    CHK(ypass(s, true, V));
    switch (V) {
      case 1: LAUNCH((k_xcontract_r<N, 1>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      case 2: LAUNCH((k_xcontract_r<N, 2>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      case 3: LAUNCH((k_xcontract_r<N, 3>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      case 4: LAUNCH((k_xcontract_r<N, 4>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, mk, s->twN, (const MemberPtrs*)nullptr); break;
      default: LAUNCH((k_xcontract<N>), gx, s->xthreads, s->lds_xc, s, s->kbuf, s->Bt, s->g, mk, s->twN);
    }
    CHK(ypass(s, false, V));
    CHK(zinv(s, out, V));
    return GSK_OK;
  }

  static int step_mart_ref(gsk_sim* s) {
    const gsk_config& c = s->cfg;
    CHK(elastic_R_real(s, s->fld, s->rbuf));
    RefMartK k{};
    k.a1 = (float)(0.02 * (c.temperature - c.T0));
    k.a2 = (float)c.B;
    k.a3 = (float)c.C;
    k.sc = (float)std::sqrt(std::fabs((0.02 * (c.temperature - c.T0)) / (c.B - c.C)));
    k.wG = (float)c.w_grad; k.wC = (float)c.w_chem; k.wE = (float)c.w_el; k.wD = (float)c.w_dis;
    k.wN = (float)c.noise; k.dt = (float)c.dt;
    k.gx = k.gy = k.gz = (float)(1.0 / (2.0 * c.dx) / 3.0);
    k.V = s->V;
    k.chem_form = c.chem_form;
    k.step = (uint32_t)s->step;
    k.stepp = s->dstep;
    k.seed = c.seed;
    CHK(launch_ref_mart(s, s->fld, s->rbuf, s->n, s->n, s->n, k));
    LAUNCH(k_advance<>, 1, 64, 0, s, s->dstep);
    // the new eta is in rbuf: the two buffers trade places instead of copying V fields back
    std::swap(s->fld, s->rbuf);
    return GSK_OK;
  }

  static int step_prec_ref(gsk_sim* s) {
    CHK(elastic_R_real(s, s->fld + nvox(s), s->rbuf));
    if (s->split) {  // local forces with R, x update, inverse (advancing the step counter)
      const PrecK pk = prec_consts(s);
      CHK(split_local(s, pk, s->rbuf, (float)(2.0 * s->cfg.xi), (float)s->cfg.noise));
      CHK(split_xupdate_prec<0>(s, pk));
      return split_inv(s, s->fld, s->F, s->dstep);
    }
    return step_prec_fused(s);
  }

  // With 2..5 fields the last z pass is fused with the next step's first one (k_zinv_local_r)
  // and the x update advances the noise counter.
  static int step_prec_fused(gsk_sim* s) {
    if (s->split) return step_prec_split(s);
    const int F = s->F;
    PrecK pk = prec_consts(s);
    const bool fuse = prec_carries(s);
    if (!fuse || !s->sq_ready) CHK(zfwd(s, Z_PREC_LOCAL, s->fld, F));
    CHK(ypass(s, true, F));
    if (fuse) pk.adv = s->dstep;
    CHK(xupdate_prec<0>(s, pk));
    CHK(ypass(s, false, F));
    if (!fuse) return zinv(s, s->fld, F, s->dstep);
    pk.adv = nullptr;
    const int rows = s->n * s->n / s->g.L;
    switch (F) {
      case 2: LAUNCH((k_zinv_local_r<M, 2>), rows, s->zthreads, zlds(s, 1), s, s->kbuf, s->fld, s->g, pk, s->twM, s->rtn); break;
      case 3: LAUNCH((k_zinv_local_r<M, 3>), rows, s->zthreads, zlds(s, 1), s, s->kbuf, s->fld, s->g, pk, s->twM, s->rtn); break;
      case 4: LAUNCH((k_zinv_local_r<M, 4>), rows, s->zthreads, zlds(s, 1), s, s->kbuf, s->fld, s->g, pk, s->twM, s->rtn); break;
      default: LAUNCH((k_zinv_local_r<M, 5>), rows, s->zthreads, zlds(s, 1), s, s->kbuf, s->fld, s->g, pk, s->twM, s->rtn);
    }
    s->sq_ready = true;
    return GSK_OK;
  }

  // ---- plain FP64 3D C2C on V fields of buf
  static int c2c(gsk_sim* s, double2* buf, bool fwd, int nf) {
    const long n = s->n;
    const long N3 = n * n * n;
    for (int f = 0; f < nf; ++f) {
      double2* b = buf + f * N3;
      const int gz = (int)(n * n / s->pL);
      if (fwd) LAUNCH((p_zrows<double2, N, true>), gz, s->pzthreads, s->lds_pz, s, b, s->pL, s->twN64);
      else LAUNCH((p_zrows<double2, N, false>), gz, s->pzthreads, s->lds_pz, s, b, s->pL, s->twN64);
      const int nch = (int)(n / s->pW), gs = (int)(n * nch);
      // y: outer x (stride n^2), positions y (stride n)
      if (fwd) LAUNCH((p_strided<double2, N, true>), gs, s->pxthreads, s->lds_px, s, b, s->pW, nch, n * n, n, s->twN64);
      else LAUNCH((p_strided<double2, N, false>), gs, s->pxthreads, s->lds_px, s, b, s->pW, nch, n * n, n, s->twN64);
      // x: outer y (stride n), positions x (stride n^2)
      if (fwd) LAUNCH((p_strided<double2, N, true>), gs, s->pxthreads, s->lds_px, s, b, s->pW, nch, n, n * n, s->twN64);
      else LAUNCH((p_strided<double2, N, false>), gs, s->pxthreads, s->lds_px, s, b, s->pW, nch, n, n * n, s->twN64);
    }
    return GSK_OK;
  }

  // Every field the step transforms is real, so two share one C2C transform and the consumer takes
  // them apart from G(k) and G(-k): squares and R_p in ceil(V/2) pairs, each force with its own
  // field (F_p + i eta_p), the new fields in pairs for the inverse (the box steps do the same).
  static int step_mart_plain(gsk_sim* s) {
    const long N3 = (long)nvox(s);
    const int P = (s->V + 1) / 2;
    PlainK k = plain_consts(s);
    LAUNCH(p_square_pair<double>, grid_1d(N3 / 4), 256, 0, s, s->eta64, s->S64, N3, s->V);
    CHK(c2c(s, s->S64, true, P));
    switch (s->V) {
#define GSK_PLAIN_CONTRACT(V) \
  case V: LAUNCH((p_contract_pair<double, V>), grid_1d(N3), 256, 0, s, s->S64, s->R64, s->B64, N3, k); break;
      GSK_PLAIN_CONTRACT(1) GSK_PLAIN_CONTRACT(2) GSK_PLAIN_CONTRACT(3) GSK_PLAIN_CONTRACT(4)
      GSK_PLAIN_CONTRACT(5) GSK_PLAIN_CONTRACT(6) GSK_PLAIN_CONTRACT(7) GSK_PLAIN_CONTRACT(8)
#undef GSK_PLAIN_CONTRACT
      default: return fail(GSK_ERR_ARG, "nvar out of range");
    }
    CHK(c2c(s, s->R64, false, P));
    LAUNCH(p_force_pair<double>, grid_1d(N3 / 4), 256, 0, s, s->eta64, s->R64, s->F64, N3, k);
    CHK(c2c(s, s->F64, true, s->V));
    LAUNCH(p_update_pair<double>, grid_1d(N3), 256, 0, s, s->F64, s->E64, k);
    CHK(c2c(s, s->E64, false, P));
    LAUNCH(p_extract_pair<double>, grid_1d(N3 / 4), 256, 0, s, s->E64, s->eta64, N3, s->V);
    return GSK_OK;
  }

  // Precipitate spectral rule on the FP64 plain path (src/dynamics_precipitate.cu:118-275): local
  // forces, C2C of forces and fields, the semi-implicit CH and AC updates, inverse C2C; forces and
  // fields share transforms (F_f + i field_f) and the new fields share theirs in pairs.
  static int step_prec_plain(gsk_sim* s) {
    const int F = s->F;
    const long N3 = (long)nvox(s);
    const PlainPK k = plain_prec_consts(s);
    LAUNCH(p_prec_local<double>, grid_1d(N3 / 4), 256, 0, s, s->eta64, s->F64, N3, k, nullptr);
    CHK(c2c(s, s->F64, true, F));
    LAUNCH(p_prec_update_pair<double>, grid_1d(N3), 256, 0, s, s->F64, s->E64, s->B64, k);
    CHK(c2c(s, s->E64, false, (F + 1) / 2));
    LAUNCH(p_extract_pair<double>, grid_1d(N3 / 4), 256, 0, s, s->E64, s->eta64, N3, F);
    return GSK_OK;
  }

  // out = {sum k^2 |c_hat|^2, sum k^2 sum_v |eta_hat_v|^2, sum B |c_hat|^2} of the plain precipitate
  static int energy_prec_plain(gsk_sim* s, double* out) {
    const long N3 = (long)nvox(s), FN = s->F * N3;
    const PlainPK k = plain_prec_consts(s);
    LAUNCH(p_load<double>, grid_1d(FN), 256, 0, s, s->eta64, s->E64, FN);
    CHK(c2c(s, s->E64, true, s->F));
    const int nb = 256;
    LAUNCH(r_plain_kspace_prec<double>, nb, 256, 0, s, s->E64, s->B64, k, s->part);
    LAUNCH(r_finish<>, 1, 1024, 0, s, s->part, nb, 3, s->rout);
    HIPCHK(hipMemcpyAsync(out, s->rout, 3 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return GSK_OK;
  }

  static int energy_mart_el_fused(gsk_sim* s, double* el) {
    if (s->split) {  // E-form S_hat; the split contraction kernel sums the energy per tile
      // and only reads the buffer, so the carried S_hat stays what the next step starts from
      if (!s->sq_ready) CHK(square_fwd(s));
      MartK mk = mart_consts(s);
      CHK(split_xcontract<1>(s, mk));
      LAUNCH(r_finish<>, 1, 1024, 0, s, s->part, s->sg.ntile_c, 1, s->rout);
      HIPCHK(hipMemcpyAsync(el, s->rout, sizeof(double), hipMemcpyDeviceToHost, s->stream));
      HIPCHK(hipStreamSynchronize(s->stream));
      return GSK_OK;
    }
    // S_hat = FFT(eta^2) through z and y; the energy kernel applies x and contracts.
    if (plane_on(s)) {
      if (!s->sq_ready) CHK(square_fwd(s));  // y,z-transformed already
    } else {
      if (!s->sq_ready) CHK(zfwd(s, Z_SQUARE, s->fld, s->F));
      CHK(ypass(s, true, s->F));
    }
    s->sq_ready = false;  // the buffer is no longer what a step starts from
    const int gx = s->g.ntile;
    if (gx > s->rblocks) return fail(GSK_ERR, "reduction scratch too small");
    switch (s->F) {
      case 1: LAUNCH((r_xenergy_mart_r<N, 1>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, s->twN, s->part); break;
      case 2: LAUNCH((r_xenergy_mart_r<N, 2>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, s->twN, s->part); break;
      case 3: LAUNCH((r_xenergy_mart_r<N, 3>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, s->twN, s->part); break;
      case 4: LAUNCH((r_xenergy_mart_r<N, 4>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->Bt, s->g, s->twN, s->part); break;
      default: {
        const size_t lds = (size_t)(N + (size_t)s->F * N * s->g.W) * sizeof(float2);
        if (lds > 160 * 1024) return fail(GSK_ERR_UNSUPPORTED, "elastic energy: too many variants for this grid");
        LAUNCH((r_xenergy_mart<N>), gx, s->xthreads, lds, s, s->kbuf, s->Bt, s->g, s->F, s->twN, s->part);
      }
    }
    LAUNCH(r_finish<>, 1, 1024, 0, s, s->part, gx, 1, s->rout);
    HIPCHK(hipMemcpyAsync(el, s->rout, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return GSK_OK;
  }

  static int energy_mart_el_plain(gsk_sim* s, double* gr, double* el) {
    const long N3 = (long)nvox(s), VN = s->V * N3;
    PlainK k = plain_consts(s);
    LAUNCH(p_square<double>, grid_1d(VN), 256, 0, s, s->eta64, s->S64, VN);
    CHK(c2c(s, s->S64, true, s->V));
    LAUNCH(p_load<double>, grid_1d(VN), 256, 0, s, s->eta64, s->E64, VN);
    CHK(c2c(s, s->E64, true, s->V));
    const int nb = 256;
    LAUNCH(r_plain_kspace<double>, nb, 256, 0, s, s->E64, s->S64, s->B64, k, s->part);
    LAUNCH(r_finish<>, 1, 1024, 0, s, s->part, nb, 2, s->rout);
    double h[2];
    HIPCHK(hipMemcpyAsync(h, s->rout, 2 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    *gr = h[0];
    *el = h[1];
    return GSK_OK;
  }
};

}  // namespace

template <int N> int impl_set_attrs() { return Impl<N>::set_attrs(); }
template <int N> int impl_init_state(gsk_sim* s) { return Impl<N>::init_state(s); }
template <int N> int impl_square_fwd(gsk_sim* s) { return Impl<N>::square_fwd(s); }
template <int N> int impl_step_batched(gsk_sim* s0, int cnt, const MemberPtrs* tab, hipStream_t st) {
  return Impl<N>::step_mart_batched(s0, cnt, tab, st);
}
template <int N> int impl_step(gsk_sim* s) {
  if (s->cfg.path == GSK_PATH_PLAIN64)
    return s->cfg.physics == GSK_PHYS_MART ? Impl<N>::step_mart_plain(s) : Impl<N>::step_prec_plain(s);
  if (s->cfg.rule == GSK_RULE_REFERENCE)
    return s->cfg.physics == GSK_PHYS_MART ? Impl<N>::step_mart_ref(s) : Impl<N>::step_prec_ref(s);
  if (s->cfg.physics == GSK_PHYS_MART) return Impl<N>::step_mart_fused(s);
  return Impl<N>::step_prec_fused(s);
}
template <int N> int impl_energy_el_fused(gsk_sim* s, double* el) { return Impl<N>::energy_mart_el_fused(s, el); }
template <int N> int impl_energy_el_plain(gsk_sim* s, double* gr, double* el) {
  return Impl<N>::energy_mart_el_plain(s, gr, el);
}
template <int N> int impl_energy_prec_plain(gsk_sim* s, double* out) { return Impl<N>::energy_prec_plain(s, out); }
template <int N> int impl_zfwd(gsk_sim* s, const float* in, int F) { return Impl<N>::zfwd(s, Z_COPY, in, F); }
template <int N> int impl_ypass(gsk_sim* s, bool fwd, int F) { return Impl<N>::ypass(s, fwd, F); }
template <int N> int impl_xpass(gsk_sim* s, bool fwd, int F, float scale) {
  const int gx = s->g.ntile;
  if (fwd) LAUNCH((k_xpass<N, true>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->g, F, scale, s->twN);
  else LAUNCH((k_xpass<N, false>), gx, s->xthreads, s->lds_y, s, s->kbuf, s->g, F, scale, s->twN);
  return GSK_OK;
}
template <int N> int impl_adj_step(gsk_sim* s, const float* eta, const float* eta1, float* lam, float* mu, float* rq,
                                   float2* kb, float2* k2, const float* bh, double* part, int nb, const AdjK& k) {
  return Impl<N>::adj_step(s, eta, eta1, lam, mu, rq, kb, k2, bh, part, nb, k);
}
template <int N> int impl_r2c(gsk_sim* s, const float* in, float2* kb, int F) { return Impl<N>::r2c_buf(s, in, kb, F); }
template <int N> int impl_c2r(gsk_sim* s, float2* kb, float* out, int F) { return Impl<N>::c2r_buf(s, kb, out, F); }
template <int N> int impl_zinv(gsk_sim* s, float* out, int F) { return Impl<N>::zinv(s, out, F); }

#define GSK_INST(N)                                                                                              \
  template int impl_set_attrs<N>();                                                                              \
  template int impl_init_state<N>(gsk_sim*);                                                                     \
  template int impl_square_fwd<N>(gsk_sim*);                                                                     \
  template int impl_step_batched<N>(gsk_sim*, int, const MemberPtrs*, hipStream_t);                              \
  template int impl_step<N>(gsk_sim*);                                                                           \
  template int impl_energy_el_fused<N>(gsk_sim*, double*);                                                       \
  template int impl_energy_el_plain<N>(gsk_sim*, double*, double*);                                              \
  template int impl_energy_prec_plain<N>(gsk_sim*, double*);                                                     \
  template int impl_zfwd<N>(gsk_sim*, const float*, int);                                                        \
  template int impl_ypass<N>(gsk_sim*, bool, int);                                                               \
  template int impl_xpass<N>(gsk_sim*, bool, int, float);                                                        \
  template int impl_adj_step<N>(gsk_sim*, const float*, const float*, float*, float*, float*, float2*, float2*,    \
                                const float*, double*, int, const AdjK&);                                         \
  template int impl_r2c<N>(gsk_sim*, const float*, float2*, int);                                                \
  template int impl_c2r<N>(gsk_sim*, float2*, float*, int);                                                      \
  template int impl_zinv<N>(gsk_sim*, float*, int);
GSK_INST(GSK_IMPL_N)
