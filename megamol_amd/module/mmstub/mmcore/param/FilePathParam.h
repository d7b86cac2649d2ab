#pragma once
#include <filesystem>
#include "mmcore/param/ParamSlot.h"
namespace megamol::core::param {
class FilePathParam : public AbstractParam { public: explicit FilePathParam(const std::string& v) : v_(v) {} std::filesystem::path Value() const { return v_; } void SetValue(const std::string& v) { v_ = v; }
    bool ParseValue(const std::string& s) override { v_ = s; return true; }
private: std::filesystem::path v_; };
}
