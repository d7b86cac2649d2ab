#pragma once
namespace megamol::core {
class Call { public: virtual ~Call() = default; };
class CallerSlot {
public:
    CallerSlot(const char* name, const char* desc) { (void)name; (void)desc; }
    template <class D> void SetCompatibleCall() {}
    template <class T> T* CallAs() { return dynamic_cast<T*>(call_); }
    void ConnectCall(Call* c) { call_ = c; }
private:
    Call* call_ = nullptr;
};
}  // namespace megamol::core
