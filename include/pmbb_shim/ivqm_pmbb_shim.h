// ivqm_pmbb_shim.h - puts the drop-in processors (ivqm_dropin.hpp) into the
// QMIV app's namespace under the reference's class names, for building the
// app's own sources against the MI355X library instead of pmbb-ivqm.
//
// Use: put this directory first on the include path, then the reference's
// pmbb-ivqm, pmbb-core and pmbb-base source directories, and link
// libivqm_hip.so in place of the pmbb-ivqm library. The app's includes of
// xIVPSNR.h, xIVSSIM.h and xShftCompPic.h (apps/QMIV/src/xAppQMIV.h:9-16) then
// resolve to the headers beside this one, which include only this file; the
// rest of pmbb-ivqm the app reads (xCommonDefIVQM.h, via xMetricQMIV.h:6) is
// the reference's own.
//
// Besides the processors this provides the parameter defaults the app reads
// (xCorrespPixelShiftPrms.h:20-22, xGlobClrDiff.h:20) with the reference's
// vector types, and the reference's xPicP to PicTraits (getAddr takes eCmp).
#pragma once
#include "ivqm_dropin.hpp"
#include "xCommonDefIVQM.h"
#include "xPic.h"
#include "xPlane.h"
#include "xVec.h"

template <> struct ivqm::dropin::PicTraits<PMBB_NAMESPACE::xPicP> {
  static constexpr bool interleaved = false;
  static const uint16_t* plane(const PMBB_NAMESPACE::xPicP* p, int c) { return p->getAddr((PMBB_NAMESPACE::eCmp)c); }
  static uint16_t* plane_mut(PMBB_NAMESPACE::xPicP* p, int c) { return p->getAddr((PMBB_NAMESPACE::eCmp)c); }
};
static_assert(ivqm::dropin::PicTraits<PMBB_NAMESPACE::xPicI>::interleaved, "xPicI::getAddr() takes no component");

// The app warns when the search has no fast path for the default component weights
// (xAppQMIV.cpp:507; xCorrespPixelShift.h:17-20 defines this per CPU): the device search has
// fast variants for any weights.
#ifndef X_CORRESPPIXELSHIFT_CAN_USE_SSE
#define X_CORRESPPIXELSHIFT_CAN_USE_SSE 1
#endif

namespace PMBB_NAMESPACE {

struct xCorrespPixelShiftPrms {
  static constexpr int32 c_DefaultSearchRange = 2;
  static constexpr int32V4 c_DefaultCmpWeights = {4, 1, 1, 0};
  static constexpr int32V4 c_EqualCmpWeights = {1, 1, 1, 0};
};
struct xGlobClrDiffPrms {
  static constexpr flt32V4 c_DefaultUnntcbCoef = {0.01f, 0.01f, 0.01f, 0.0f};
};

using xGlobClrDiffProc = ivqm::dropin::xGlobClrDiffProc;
using xShftCompPicProc = ivqm::dropin::xShftCompPicProc;
using xPSNR = ivqm::dropin::xPSNR;
using xWSPSNR = ivqm::dropin::xWSPSNR;
using xIVPSNR = ivqm::dropin::xIVPSNR;
using xIVPSNRM = ivqm::dropin::xIVPSNRM;
using xSSIM = ivqm::dropin::xSSIM;
using xIVSSIM = ivqm::dropin::xIVSSIM;

}  // namespace PMBB_NAMESPACE
