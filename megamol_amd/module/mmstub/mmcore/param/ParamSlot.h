#pragma once
#include <memory>
#include <string>
namespace megamol::core::param {
// AbstractParam::ParseValue (core/include/mmcore/param/AbstractParam.h): a value from its string
// form, as project files (mmSetParamValue) set it
class AbstractParam { public: virtual ~AbstractParam() = default; virtual bool ParseValue(const std::string& v) = 0; };
class ParamSlot {
public:
    ParamSlot(const char* name, const char* desc) : name_(name) { (void)desc; }
    ParamSlot& operator<<(AbstractParam* p) { p_.reset(p); return *this; }
    template <class P> P* Param() { return dynamic_cast<P*>(p_.get()); }
    AbstractParam* Parameter() { return p_.get(); }
    const std::string& Name() const { return name_; }
private:
    std::string name_;
    std::unique_ptr<AbstractParam> p_;
};
}  // namespace megamol::core::param
