/*
 * C++ adapter with the shape of the reference's per-box operator,
 *   BoxOp_EulerCubedSphere<T,MAP,MEM>::operator()(a_Rhs, a_fluxes, a_USph, a_dt,
 *       a_scale, a_counter, a_cartesianPrimitives, a_activeFaces)
 *   (src/BoxOp_EulerCubedSphere.H:412-421, src/BoxOp_EulerCubedSphereImplem.H:4-402)
 * and BertaOperator (src/BoxOp_EulerCubedSphere.H:423-430), over the C ABI of
 * heliocubed.h: one operator per box of the layout, as the reference builds
 * one per box, evaluated on the device. Header only, C++17; links against
 * libheliocubed_amd.so.
 *
 * Arguments keep the reference's meaning:
 *   a_Rhs     the valid box; the mapped flux divergence times a_scale is ADDED
 *             (a_Rhs += rhsIncr, src/BoxOp_EulerCubedSphereImplem.H:355-378);
 *   a_fluxes  face fluxes of the valid box extruded in each direction (a view
 *             with data == nullptr is not written);
 *   a_USph    at fourth order the native conserved cell averages on the valid
 *             box grown by ghost() (5); at second order the Cartesian
 *             primitives on the box grown by 2 (a_cartesianPrimitives true,
 *             the StageLevel call, src/BoxOp_EulerCubedSphereImplem.H:1003-1004;
 *             the native-conserved form is not built and is rejected);
 *   a_dt, a_counter  accepted and unused, as in the reference;
 *   a_activeFaces  the composite-AMR face mask of the second-order operator:
 *             not built (nullptr required), ignored at fourth order.
 * spatial_order 2 selects secondOrderOperator (src/BoxOp_SecondOrder.H:155-370);
 * at order 4 the options select the standard operator or, with
 * spatial_operator 2, the Berta operator, which operator() calls as the
 * reference does (src/BoxOp_EulerCubedSphereImplem.H:73-86).
 * Errors (a view whose box does not match, a failed device call) throw
 * std::runtime_error with the library's message: the reference aborts
 * through MayDay on the same conditions.
 */
#ifndef HELIOCUBED_BOXOP_HPP
#define HELIOCUBED_BOXOP_HPP

#include <array>
#include <stdexcept>
#include <string>

#include "heliocubed.h"

namespace heliocubed
{

/* Device view of a Proto BoxData<double,8> on the box [lo, hi]: data points
 * at component 0 of cell lo, components outermost, radial index fastest
 * ([8][n2][n1][n0] doubles). */
struct DeviceBoxData
{
    double* data = nullptr;
    int lo[3] = {0, 0, 0};
    int hi[3] = {-1, -1, -1};
    long long size(int d) const { return (long long)hi[d] - lo[d] + 1; }
};

class BoxOp_EulerCubedSphere
{
public:
    /* The operator of the valid box [lo, hi] of panel a_block, on the device
     * and stream current at construction (stream may be nullptr). */
    BoxOp_EulerCubedSphere(const hc_shell& shell, const hc_options& options, unsigned block, const int lo[3],
                           const int hi[3], hc_stream stream = nullptr)
        : m_options(options), m_block(block), m_stream(stream)
    {
        hc_box box{(int)block, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}};
        for (int d = 0; d < 3; ++d) { m_lo[d] = lo[d]; m_hi[d] = hi[d]; }
        const int rc = options.spatial_order == 2 ? hc_boxop2_create(&shell, &options, &box, 1, &m_op2)
                                                  : hc_boxop_create(&shell, &options, &box, 1, &m_op);
        if (rc != 0) throw std::runtime_error(hc_last_error());
    }
    ~BoxOp_EulerCubedSphere() { release(); }
    BoxOp_EulerCubedSphere(const BoxOp_EulerCubedSphere&) = delete;
    BoxOp_EulerCubedSphere& operator=(const BoxOp_EulerCubedSphere&) = delete;
    BoxOp_EulerCubedSphere(BoxOp_EulerCubedSphere&& o) noexcept { *this = static_cast<BoxOp_EulerCubedSphere&&>(o); }
    BoxOp_EulerCubedSphere& operator=(BoxOp_EulerCubedSphere&& o) noexcept
    {
        if (this != &o)
        {
            release();
            m_op = o.m_op; o.m_op = nullptr;
            m_op2 = o.m_op2; o.m_op2 = nullptr;
            m_options = o.m_options; m_block = o.m_block; m_stream = o.m_stream;
            for (int d = 0; d < 3; ++d) { m_lo[d] = o.m_lo[d]; m_hi[d] = o.m_hi[d]; }
        }
        return *this;
    }

    /* src/BoxOp_EulerCubedSphere.H:376-396: ghost 2 at second order, 5 at
     * fourth order (the Jameson detector of the Berta operator) */
    static constexpr int numState() { return 8; }
    int ghost() const { return m_options.spatial_order == 2 ? 2 : 5; }
    int order() const { return m_options.spatial_order == 2 ? 2 : 4; }
    unsigned block() const { return m_block; }
    void setStream(hc_stream stream) { m_stream = stream; }

    void operator()(DeviceBoxData& a_Rhs, std::array<DeviceBoxData, 3>& a_fluxes, const DeviceBoxData& a_USph,
                    double a_dt, double a_scale = 1.0, int a_counter = -1, bool a_cartesianPrimitives = false,
                    const void* a_activeFaces = nullptr) const
    {
        if (m_options.spatial_order == 2)
        {
            if (!a_cartesianPrimitives)
                throw std::runtime_error("BoxOp_EulerCubedSphere: the second-order operator takes Cartesian primitives "
                                         "(a_cartesianPrimitives true)");
            if (a_activeFaces)
                throw std::runtime_error("BoxOp_EulerCubedSphere: the composite-AMR face mask is not built");
            apply(a_Rhs, a_fluxes, a_USph, a_scale);
            return;
        }
        if (m_options.spatial_operator == 2) { BertaOperator(a_Rhs, a_fluxes, a_USph, a_dt, a_scale, a_counter); return; }
        apply(a_Rhs, a_fluxes, a_USph, a_scale);
    }

    void BertaOperator(DeviceBoxData& a_Rhs, std::array<DeviceBoxData, 3>& a_fluxes, const DeviceBoxData& a_USph,
                       double a_dt, double a_scale = 1.0, int a_counter = -1) const
    {
        (void)a_dt; (void)a_counter;
        if (m_options.spatial_order != 4 || m_options.spatial_operator != 2)
            throw std::runtime_error("BoxOp_EulerCubedSphere::BertaOperator: the operator was built with spatial_operator "
                                     + std::to_string(m_options.spatial_operator));
        apply(a_Rhs, a_fluxes, a_USph, a_scale);
    }

private:
    void expect(const DeviceBoxData& b, const int lo[3], const int hi[3], const char* what) const
    {
        for (int d = 0; d < 3; ++d)
            if (b.lo[d] != lo[d] || b.hi[d] != hi[d])
                throw std::runtime_error(std::string("BoxOp_EulerCubedSphere: ") + what + " is not on the expected box");
    }
    void apply(DeviceBoxData& rhs, std::array<DeviceBoxData, 3>& fluxes, const DeviceBoxData& usph, double scale) const
    {
        if (!rhs.data || !usph.data) throw std::runtime_error("BoxOp_EulerCubedSphere: null Rhs or USph data");
        int glo[3], ghi[3];
        for (int d = 0; d < 3; ++d) { glo[d] = m_lo[d] - ghost(); ghi[d] = m_hi[d] + ghost(); }
        expect(usph, glo, ghi, "USph");
        expect(rhs, m_lo, m_hi, "Rhs");
        double* F[3] = {nullptr, nullptr, nullptr};
        for (int d = 0; d < 3; ++d)
        {
            if (!fluxes[d].data) continue;
            int fhi[3] = {m_hi[0], m_hi[1], m_hi[2]};
            fhi[d] += 1;
            expect(fluxes[d], m_lo, fhi, "a flux");
            F[d] = fluxes[d].data;
        }
        const int rc = m_op2 ? hc_boxop2_apply(m_op2, usph.data, rhs.data, F[0], F[1], F[2], scale, m_stream)
                             : hc_boxop_apply_scaled(m_op, usph.data, rhs.data, F[0], F[1], F[2], scale, m_stream);
        if (rc != 0) throw std::runtime_error(hc_last_error());
    }
    void release()
    {
        if (m_op) hc_boxop_destroy(m_op);
        if (m_op2) hc_boxop2_destroy(m_op2);
        m_op = nullptr; m_op2 = nullptr;
    }

    hc_boxop* m_op = nullptr;
    hc_boxop2* m_op2 = nullptr;
    hc_options m_options{};
    unsigned m_block = 0;
    int m_lo[3] = {0, 0, 0}, m_hi[3] = {-1, -1, -1};
    hc_stream m_stream = nullptr;
};

} // namespace heliocubed

#endif
