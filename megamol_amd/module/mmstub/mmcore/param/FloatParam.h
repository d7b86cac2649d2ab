#pragma once
#include <cstdlib>
#include "mmcore/param/ParamSlot.h"
namespace megamol::core::param {
class FloatParam : public AbstractParam { public: FloatParam(float v, float mn = -3.4e38f) : v_(v) { (void)mn; } float Value() const { return v_; } void SetValue(float v) { v_ = v; }
    bool ParseValue(const std::string& s) override { char* e = nullptr; float v = std::strtof(s.c_str(), &e); if (!e || *e) return false; v_ = v; return true; }
private: float v_; };
}
