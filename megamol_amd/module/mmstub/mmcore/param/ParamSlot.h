#pragma once
#include <memory>
#include <string>
namespace megamol::core::param {
class AbstractParam { public: virtual ~AbstractParam() = default; };
class ParamSlot {
public:
    ParamSlot(const char* name, const char* desc) : name_(name) { (void)desc; }
    ParamSlot& operator<<(AbstractParam* p) { p_.reset(p); return *this; }
    template <class P> P* Param() { return dynamic_cast<P*>(p_.get()); }
    const std::string& Name() const { return name_; }
private:
    std::string name_;
    std::unique_ptr<AbstractParam> p_;
};
}  // namespace megamol::core::param
