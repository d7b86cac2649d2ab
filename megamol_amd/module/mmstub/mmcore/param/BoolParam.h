#pragma once
#include "mmcore/param/ParamSlot.h"
namespace megamol::core::param {
class BoolParam : public AbstractParam { public: explicit BoolParam(bool v) : v_(v) {} bool Value() const { return v_; } void SetValue(bool v) { v_ = v; }
    bool ParseValue(const std::string& s) override { if (s == "true" || s == "1") v_ = true; else if (s == "false" || s == "0") v_ = false; else return false; return true; }
private: bool v_; };
}
