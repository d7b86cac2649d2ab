#pragma once
#include <map>
#include <string>
#include "mmcore/param/ParamSlot.h"
namespace megamol::core::param {
class EnumParam : public AbstractParam { public: explicit EnumParam(int v) : v_(v) {} void SetTypePair(int v, const char* n) { names_[v] = n; } int Value() const { return v_; } void SetValue(int v) { v_ = v; } private: int v_; std::map<int, std::string> names_; };
}
