#pragma once
#include "mmcore/param/ParamSlot.h"
namespace megamol::core::param {
class FloatParam : public AbstractParam { public: FloatParam(float v, float mn = -3.4e38f) : v_(v) { (void)mn; } float Value() const { return v_; } void SetValue(float v) { v_ = v; } private: float v_; };
}
