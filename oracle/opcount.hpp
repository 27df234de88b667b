// Operation-counting scalar for the oracle (test infrastructure): a double
// with the same layout whose arithmetic increments per-thread counters, so
// the oracle compiled with it (hc_opcount.cpp) reports the floating-point
// operations of its box operators. Geometry helpers switch counting off in
// their scope (ORC_UNCOUNTED), so only the per-cell and per-face arithmetic
// of the operator is counted, as the device path computes its metric tables
// once.
#pragma once
#include <cmath>
#include <limits>
#include <type_traits>

struct OpCounts
{
    long long add = 0, mul = 0, div = 0, sqrt = 0, cmp = 0, other = 0;
};

inline OpCounts& opCounts()
{
    static thread_local OpCounts c;
    return c;
}
inline bool& opCounting()
{
    static thread_local bool on = true;
    return on;
}

// Counting off in a scope.
struct OpPause
{
    bool saved;
    OpPause() : saved(opCounting()) { opCounting() = false; }
    ~OpPause() { opCounting() = saved; }
};

struct Counted
{
    double v;
    constexpr Counted() : v(0.0) {}
    constexpr Counted(double x) : v(x) {}
    template <typename I, typename = std::enable_if_t<std::is_integral<I>::value>>
    constexpr Counted(I x) : v(double(x)) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return float(v); }
    explicit operator int() const { return int(v); }
    explicit operator long() const { return long(v); }
    explicit operator long long() const { return (long long)v; }
    explicit operator unsigned long() const { return (unsigned long)v; }
    explicit operator bool() const { return v != 0.0; }
    Counted& operator+=(Counted o);
    Counted& operator-=(Counted o);
    Counted& operator*=(Counted o);
    Counted& operator/=(Counted o);
};
static_assert(sizeof(Counted) == sizeof(double), "Counted must alias double arrays");

#define OPC_TICK(field) do { if (opCounting()) ++opCounts().field; } while (0)

inline Counted operator+(Counted a, Counted b) { OPC_TICK(add); return Counted(a.v + b.v); }
inline Counted operator-(Counted a, Counted b) { OPC_TICK(add); return Counted(a.v - b.v); }
inline Counted operator*(Counted a, Counted b) { OPC_TICK(mul); return Counted(a.v * b.v); }
inline Counted operator/(Counted a, Counted b) { OPC_TICK(div); return Counted(a.v / b.v); }
inline Counted operator-(Counted a) { return Counted(-a.v); }
inline Counted operator+(Counted a) { return a; }
inline Counted& Counted::operator+=(Counted o) { *this = *this + o; return *this; }
inline Counted& Counted::operator-=(Counted o) { *this = *this - o; return *this; }
inline Counted& Counted::operator*=(Counted o) { *this = *this * o; return *this; }
inline Counted& Counted::operator/=(Counted o) { *this = *this / o; return *this; }

#define OPC_CMP(op) \
    inline bool operator op(Counted a, Counted b) { OPC_TICK(cmp); return a.v op b.v; }
OPC_CMP(<) OPC_CMP(>) OPC_CMP(<=) OPC_CMP(>=) OPC_CMP(==) OPC_CMP(!=)
#undef OPC_CMP

// mixed operands: every arithmetic type converts to Counted
#define OPC_MIX(op)                                                                                \
    template <typename T, typename = std::enable_if_t<std::is_arithmetic<T>::value>>              \
    inline auto operator op(Counted a, T b) { return a op Counted(double(b)); }                   \
    template <typename T, typename = std::enable_if_t<std::is_arithmetic<T>::value>>              \
    inline auto operator op(T a, Counted b) { return Counted(double(a)) op b; }
OPC_MIX(+) OPC_MIX(-) OPC_MIX(*) OPC_MIX(/) OPC_MIX(<) OPC_MIX(>) OPC_MIX(<=) OPC_MIX(>=) OPC_MIX(==) OPC_MIX(!=)
#undef OPC_MIX

namespace std
{
inline Counted sqrt(Counted a) { OPC_TICK(sqrt); return Counted(std::sqrt(a.v)); }
inline Counted abs(Counted a) { OPC_TICK(other); return Counted(std::fabs(a.v)); }
inline Counted fabs(Counted a) { OPC_TICK(other); return Counted(std::fabs(a.v)); }
inline Counted floor(Counted a) { OPC_TICK(other); return Counted(std::floor(a.v)); }
inline Counted round(Counted a) { OPC_TICK(other); return Counted(std::round(a.v)); }
inline long lround(Counted a) { OPC_TICK(other); return std::lround(a.v); }
inline Counted pow(Counted a, Counted b) { OPC_TICK(other); return Counted(std::pow(a.v, b.v)); }
inline Counted pow(Counted a, int b) { OPC_TICK(other); return Counted(std::pow(a.v, b)); }
inline Counted pow(Counted a, double b) { OPC_TICK(other); return Counted(std::pow(a.v, b)); }
inline Counted exp(Counted a) { OPC_TICK(other); return Counted(std::exp(a.v)); }
inline Counted expm1(Counted a) { OPC_TICK(other); return Counted(std::expm1(a.v)); }
inline Counted log(Counted a) { OPC_TICK(other); return Counted(std::log(a.v)); }
inline Counted sin(Counted a) { OPC_TICK(other); return Counted(std::sin(a.v)); }
inline Counted cos(Counted a) { OPC_TICK(other); return Counted(std::cos(a.v)); }
inline Counted tan(Counted a) { OPC_TICK(other); return Counted(std::tan(a.v)); }
inline Counted atan(Counted a) { OPC_TICK(other); return Counted(std::atan(a.v)); }
inline Counted atan2(Counted a, Counted b) { OPC_TICK(other); return Counted(std::atan2(a.v, b.v)); }
inline Counted acos(Counted a) { OPC_TICK(other); return Counted(std::acos(a.v)); }
inline Counted fmax(Counted a, Counted b) { OPC_TICK(cmp); return Counted(std::fmax(a.v, b.v)); }
inline Counted fmod(Counted a, Counted b) { OPC_TICK(other); return Counted(std::fmod(a.v, b.v)); }
inline Counted fmod(Counted a, double b) { OPC_TICK(other); return Counted(std::fmod(a.v, b)); }
inline Counted fmod(Counted a, int b) { OPC_TICK(other); return Counted(std::fmod(a.v, double(b))); }
inline Counted fmin(Counted a, Counted b) { OPC_TICK(cmp); return Counted(std::fmin(a.v, b.v)); }
inline bool isfinite(Counted a) { return std::isfinite(a.v); }
inline bool isnan(Counted a) { return std::isnan(a.v); }
// max/min with one plain-double operand (the templates need equal types)
inline const Counted max(const Counted& a, const Counted& b) { OPC_TICK(cmp); return a.v < b.v ? b : a; }
inline const Counted min(const Counted& a, const Counted& b) { OPC_TICK(cmp); return b.v < a.v ? b : a; }
inline const Counted max(const Counted& a, double b) { return std::max(a, Counted(b)); }
inline const Counted max(double a, const Counted& b) { return std::max(Counted(a), b); }
inline const Counted min(const Counted& a, double b) { return std::min(a, Counted(b)); }
inline const Counted min(double a, const Counted& b) { return std::min(Counted(a), b); }

template <>
struct numeric_limits<Counted> : numeric_limits<double>
{
    static Counted min() { return Counted(numeric_limits<double>::min()); }
    static Counted max() { return Counted(numeric_limits<double>::max()); }
    static Counted epsilon() { return Counted(numeric_limits<double>::epsilon()); }
    static Counted infinity() { return Counted(numeric_limits<double>::infinity()); }
    static Counted quiet_NaN() { return Counted(numeric_limits<double>::quiet_NaN()); }
};
} // namespace std
