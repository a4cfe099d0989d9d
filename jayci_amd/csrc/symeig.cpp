// Dense symmetric eigensolver for the Davidson subspace (Rayleigh-Ritz, Loewdin
// orthonormalisation): Householder reduction to tridiagonal form with the transformations
// accumulated, then the implicit QL iteration with Wilkinson shifts on the tridiagonal
// matrix (the EISPACK tred2 / tql2 pair, Wilkinson & Reinsch, Handbook for Automatic
// Computation vol. II), written for row-major input.  O(n^3) with a small constant: a
// 20 x 20 subspace takes microseconds, which matters when a batch step solves one per
// instance.  Output: eigenvalues ascending, eigenvector i in vec[i*n .. i*n+n), each with its
// largest-magnitude component positive (deterministic vectors).
#include <math.h>
#include <string.h>
#include <algorithm>
#include <vector>

void jci_sym_eig(int n, const double *a_in, double *eval, double *vec)
{
    if (n <= 0) return;
    if (n == 1) {
        eval[0] = a_in[0];
        vec[0] = 1.0;
        return;
    }
    // z: working matrix, row-major.  Qt holds the accumulated transformation transposed
    // (row j of Qt = column j of Q), so reflections and rotations touch contiguous rows.
    std::vector<double> z(a_in, a_in + (size_t)n * n), d(n), e(n, 0.0), v(n), p(n), w(n);
    std::vector<double> Qt((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) Qt[(size_t)i * n + i] = 1.0;
    auto A = [&](int i, int j) -> double & { return z[(size_t)i * n + j]; };
    // Householder: for column k, reflect rows/columns k+1..n-1 so that A[k+2.., k] = 0
    for (int k = 0; k < n - 2; k++) {
        double norm2 = 0.0;
        for (int i = k + 1; i < n; i++) norm2 += A(i, k) * A(i, k);
        const double xnorm = sqrt(norm2);
        if (xnorm == 0.0) continue;
        const double x0 = A(k + 1, k);
        const double alpha = x0 >= 0.0 ? -xnorm : xnorm;
        // v = x - alpha e1, normalised
        double vn2 = 0.0;
        for (int i = k + 1; i < n; i++) {
            v[i] = A(i, k) - (i == k + 1 ? alpha : 0.0);
            vn2 += v[i] * v[i];
        }
        if (vn2 == 0.0) continue;
        const double vs = 1.0 / sqrt(vn2);
        for (int i = k + 1; i < n; i++) v[i] *= vs;
        // trailing block: A <- H A H with H = I - 2 v v^T, as a rank-2 update
        // p = A v, K = v^T p, q = p - K v, A -= 2 (v q^T + q v^T)
        double K = 0.0;
        for (int i = k + 1; i < n; i++) {
            double s = 0.0;
            for (int j = k + 1; j < n; j++) s += A(i, j) * v[j];
            p[i] = s;
            K += v[i] * s;
        }
        for (int i = k + 1; i < n; i++) p[i] -= K * v[i];
        for (int i = k + 1; i < n; i++)
            for (int j = k + 1; j < n; j++) A(i, j) -= 2.0 * (v[i] * p[j] + p[i] * v[j]);
        A(k + 1, k) = A(k, k + 1) = alpha;
        for (int i = k + 2; i < n; i++) A(i, k) = A(k, i) = 0.0;
        // Q <- Q H, i.e. Qt <- H Qt on rows k+1..n-1: Qt -= 2 v (v^T Qt)
        for (int c = 0; c < n; c++) w[c] = 0.0;
        for (int j = k + 1; j < n; j++) {
            const double vj = v[j];
            const double *row = &Qt[(size_t)j * n];
            for (int c = 0; c < n; c++) w[c] += vj * row[c];
        }
        for (int j = k + 1; j < n; j++) {
            const double vj = 2.0 * v[j];
            double *row = &Qt[(size_t)j * n];
            for (int c = 0; c < n; c++) row[c] -= vj * w[c];
        }
    }
    for (int i = 0; i < n; i++) d[i] = A(i, i);
    for (int i = 0; i < n - 1; i++) e[i] = A(i + 1, i);   // e[i] couples i and i+1
    e[n - 1] = 0.0;
    // implicit QL with Wilkinson shifts; rotations applied to the columns of Q
    const double eps = 2.220446049250313e-16;
    for (int l = 0; l < n; l++) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; m++) {
                const double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (++iter > 60) break;   // no convergence: keep what we have
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
            double s = 1.0, c = 1.0, pp = 0.0;
            int i = m - 1;
            bool underflow = false;
            for (; i >= l; i--) {
                double f = s * e[i], b = c * e[i];
                r = hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= pp;
                    e[m] = 0.0;
                    underflow = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - pp;
                r = (d[i] - g) * s + 2.0 * c * b;
                pp = s * r;
                d[i + 1] = g + pp;
                g = c * r - b;
                double *r0 = &Qt[(size_t)i * n], *r1 = &Qt[(size_t)(i + 1) * n];
                for (int k = 0; k < n; k++) {
                    const double q1 = r1[k], q0 = r0[k];
                    r1[k] = s * q0 + c * q1;
                    r0[k] = c * q0 - s * q1;
                }
            }
            if (underflow) continue;
            d[l] -= pp;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    std::vector<int> ord(n);
    for (int i = 0; i < n; i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return d[x] < d[y]; });
    for (int i = 0; i < n; i++) {
        const int c = ord[i];
        const double *row = &Qt[(size_t)c * n];
        eval[i] = d[c];
        int arg = 0;
        for (int k = 1; k < n; k++)
            if (fabs(row[k]) > fabs(row[arg])) arg = k;
        const double sg = row[arg] < 0.0 ? -1.0 : 1.0;
        for (int k = 0; k < n; k++) vec[(size_t)i * n + k] = sg * row[k];
    }
}
