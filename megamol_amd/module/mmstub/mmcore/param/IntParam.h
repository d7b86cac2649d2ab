#pragma once
#include "mmcore/param/ParamSlot.h"
namespace megamol::core::param {
class IntParam : public AbstractParam { public: IntParam(int v, int mn = -2147483647, int mx = 2147483647) : v_(v) { (void)mn; (void)mx; } int Value() const { return v_; } void SetValue(int v) { v_ = v; } private: int v_; };
}
