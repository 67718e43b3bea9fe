// qft.hip - quantum Fourier transform on a contiguous qubit range.
//
// The reference builds QFT/IQFT from Hadamards, controlled phase shifts and
// a SWAP network (QuantumFourierTransform.h:35-87, QubitsSwapper.h:23-33).
// qft_sequence() issues exactly that sequence on the in-place kernels: each
// controlled phase touches only the quarter of the state with both bits set
// (the reference sweeps every amplitude per gate).
#include <cmath>
#include <complex>

#include "common.h"

namespace qcs {

int qft_sequence(qcs_register* r, uint32_t sq, uint32_t eq, int do_swap, int inverse) {
    const double hs = 1.0 / std::sqrt(2.0);
    const hcplx H[4] = {{hs, 0}, {hs, 0}, {hs, 0}, {-hs, 0}};
    auto had = [&](uint32_t q) { return k1q(r, 0, q, nullptr, 0, H); };
    auto cphase = [&](uint32_t cur, uint32_t ctrl, double phase) {
        const std::complex<double> p = std::polar(1.0, phase);
        const uint32_t bits[2] = {cur, ctrl};
        return kphase(r, bits, 2, hcplx{p.real(), p.imag()});
    };
    auto swaps = [&]() {
        for (uint32_t a = sq, b = eq; a < b; ++a, --b) {
            const int e = kswap(r, a, b, nullptr, 0, hcplx{1, 0}, false);
            if (e) return e;
        }
        return 0;
    };
    const int s = int(sq), e = int(eq);
    int err = 0;
    if (!inverse) {
        if ((err = had(eq))) return err;
        for (int cur = e; cur > s; --cur) {
            double phase = M_PI_2;
            for (int ctrl = cur - 1; ctrl >= s; --ctrl) {
                if ((err = cphase(cur, ctrl, phase))) return err;
                phase *= 0.5;
            }
            if ((err = had(cur - 1))) return err;
        }
        if (do_swap && (err = swaps())) return err;
    } else {
        if (do_swap && (err = swaps())) return err;
        for (int cur = s + 1; cur <= e; ++cur) {
            if ((err = had(cur - 1))) return err;
            double phase = -M_PI_2;
            for (int ctrl = cur - 1; ctrl >= s; --ctrl) {
                if ((err = cphase(cur, ctrl, phase))) return err;
                phase *= 0.5;
            }
        }
        if ((err = had(eq))) return err;
    }
    return QCS_OK;
}

int qft(qcs_register* r, uint32_t sq, uint32_t eq, int do_swap, int inverse) {
    return qft_sequence(r, sq, eq, do_swap, inverse);
}

}  // namespace qcs
