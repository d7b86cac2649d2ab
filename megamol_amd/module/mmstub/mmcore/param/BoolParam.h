#pragma once
#include "mmcore/param/ParamSlot.h"
namespace megamol::core::param {
class BoolParam : public AbstractParam { public: explicit BoolParam(bool v) : v_(v) {} bool Value() const { return v_; } void SetValue(bool v) { v_ = v; } private: bool v_; };
}
