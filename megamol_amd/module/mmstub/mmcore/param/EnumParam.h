#pragma once
#include <cstdlib>
#include <map>
#include <string>
#include "mmcore/param/ParamSlot.h"
namespace megamol::core::param {
class EnumParam : public AbstractParam { public: explicit EnumParam(int v) : v_(v) {} void SetTypePair(int v, const char* n) { names_[v] = n; } int Value() const { return v_; } void SetValue(int v) { v_ = v; }
    bool ParseValue(const std::string& s) override { for (auto& kv : names_) if (kv.second == s) { v_ = kv.first; return true; } char* e = nullptr; long v = std::strtol(s.c_str(), &e, 10); if (!e || *e || !names_.count((int)v)) return false; v_ = (int)v; return true; }
private: int v_; std::map<int, std::string> names_; };
}
