#pragma once
#include <string>
namespace megamol::core {
class Call { public: virtual ~Call() = default; };
class CallerSlot {
public:
    CallerSlot(const char* name, const char* desc) : name_(name) { (void)desc; }
    template <class D> void SetCompatibleCall() {}
    template <class T> T* CallAs() { return dynamic_cast<T*>(call_); }
    void ConnectCall(Call* c) { call_ = c; }
    const std::string& Name() const { return name_; }
private:
    std::string name_;
    Call* call_ = nullptr;
};
}  // namespace megamol::core
