/* pct_net_basis.h -- generated by gen_net_basis.py; do not edit.
 *
 * Fixed tables of the learned-path form, shared by the HIP kernels and the
 * CPU oracle so that both step the surrogate path with the same floats.
 * Segment i of the normalised depth sigma = u_1 / u_2 covers [i/5, (i+1)/5);
 * tau = 5 sigma - i; each row holds the coefficients c0..c3 of
 * c0 + c1 tau + c2 tau^2 + c3 tau^3.
 *   NET_H01[i]    : Hermite 3 s^2 - 2 s^3 (weight of t_out - t_in)
 *   NET_H11[i]    : Hermite s^3 - s^2     (weight of u_2 * dtheta)
 *   NET_PHI[j][i] : clamped cubic spline, 1 at knot (j+1)/5, 0 at the other
 *                   knots, zero slope at sigma = 0 and 1 (weight of offset j)
 */
#ifndef PCT_NET_BASIS_H
#define PCT_NET_BASIS_H

#define PCT_NET_SEGMENTS 5
#define PCT_NET_KNOTS 4

static constexpr float NET_H01[PCT_NET_SEGMENTS][4] = {
    {0.0f, 0.0f, 0.12f, -0.016f},
    {0.104f, 0.192f, 0.072f, -0.016f},
    {0.352f, 0.288f, 0.024f, -0.016f},
    {0.648f, 0.288f, -0.024f, -0.016f},
    {0.896f, 0.192f, -0.072f, -0.016f},
};
static constexpr float NET_H11[PCT_NET_SEGMENTS][4] = {
    {0.0f, 0.0f, -0.04f, 0.008f},
    {-0.032f, -0.056f, -0.016f, 0.008f},
    {-0.096f, -0.064f, 0.008f, 0.008f},
    {-0.144f, -0.024f, 0.032f, 0.008f},
    {-0.128f, 0.064f, 0.056f, 0.008f},
};
static constexpr float NET_PHI[PCT_NET_KNOTS][PCT_NET_SEGMENTS][4] = {
    {
        {0.0f, 0.0f, 2.784689f, -1.784689f},
        {1.0f, 0.215311005f, -2.56937799f, 1.35406699f},
        {0.0f, -0.861244019f, 1.49282297f, -0.631578947f},
        {0.0f, 0.229665072f, -0.401913876f, 0.172248804f},
        {0.0f, -0.0574162679f, 0.114832536f, -0.0574162679f},
    },
    {
        {0.0f, 0.0f, -0.746411483f, 0.746411483f},
        {0.0f, 0.746411483f, 1.49282297f, -1.23923445f},
        {1.0f, 0.014354067f, -2.22488038f, 1.21052632f},
        {0.0f, -0.803827751f, 1.40669856f, -0.602870813f},
        {0.0f, 0.200956938f, -0.401913876f, 0.200956938f},
    },
    {
        {0.0f, 0.0f, 0.200956938f, -0.200956938f},
        {0.0f, -0.200956938f, -0.401913876f, 0.602870813f},
        {0.0f, 0.803827751f, 1.40669856f, -1.21052632f},
        {1.0f, -0.014354067f, -2.22488038f, 1.23923445f},
        {0.0f, -0.746411483f, 1.49282297f, -0.746411483f},
    },
    {
        {0.0f, 0.0f, -0.0574162679f, 0.0574162679f},
        {0.0f, 0.0574162679f, 0.114832536f, -0.172248804f},
        {0.0f, -0.229665072f, -0.401913876f, 0.631578947f},
        {0.0f, 0.861244019f, 1.49282297f, -1.35406699f},
        {1.0f, -0.215311005f, -2.56937799f, 1.784689f},
    },
};

#endif
