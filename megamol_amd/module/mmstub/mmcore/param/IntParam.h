#pragma once
#include <cstdlib>
#include "mmcore/param/ParamSlot.h"
namespace megamol::core::param {
class IntParam : public AbstractParam { public: IntParam(int v, int mn = -2147483647, int mx = 2147483647) : v_(v) { (void)mn; (void)mx; } int Value() const { return v_; } void SetValue(int v) { v_ = v; }
    bool ParseValue(const std::string& s) override { char* e = nullptr; long v = std::strtol(s.c_str(), &e, 10); if (!e || *e) return false; v_ = (int)v; return true; }
private: int v_; };
}
