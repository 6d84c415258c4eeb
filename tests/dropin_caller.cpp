// dropin_caller.cpp -- a C++ caller written the way the SIP calls the
// reference's tensor module: the module's C prototypes in their by-reference
// form (tensor_ops_c_prototypes.h:41-178 declares them so), linked against
// libsipgpu.so in place of the module.  It checks every symbol against plain
// loops computed here.
//
//   dropin_caller          all checks; exit status 0 when every one passes
//   dropin_caller --no-gpu host-only symbols must work, device-backed ones
//                          must report an error (no CPU fallback)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

extern "C" {
long long int tensor_size_by_shape_(int &, int *, int &);
void get_contraction_ptrn_(int &drank, int &lrank, int &rrank, int *aces_ptrn, int *my_ptrn,
                           int &ierr);
void tensor_block_init__(int &, double *, int &, int *, double &, int &);
void tensor_block_scale__(int &, double *, int &, int *, double &, int &);
double tensor_block_norm2__(int &, double *, int &, int *, int &);
void tensor_block_slice__(int &, int &, double *, int *, double *, int *, int *, int &);
void tensor_block_insert__(int &, int &, double *, int *, double *, int *, int *, int &);
void tensor_block_add__(const int &, const int &, int *, double *, double *, const double &,
                        int &);
void tensor_block_copy__(int &, int &, int *, int *, double *, double *, int &);
void tensor_block_contract__(int &, int *, double *, int &, int *, double *, int &, int *, double *,
                             int &, int *, int &);
}

static int failures = 0;

static void check(bool ok, const char *what) {
    std::printf("%-58s %s\n", what, ok ? "ok" : "FAILED");
    if (!ok) ++failures;
}

// Deterministic test data: distinct, exactly representable values.
static std::vector<double> data(int n, int seed) {
    std::vector<double> v(n);
    unsigned s = 2654435761u * (unsigned)(seed + 1);
    for (int i = 0; i < n; ++i) {
        s = s * 1664525u + 1013904223u;
        v[i] = (double)((int)(s >> 9) % 2001 - 1000) / 64.0;
    }
    return v;
}

int main(int argc, char **argv) {
    const bool no_gpu = argc > 1 && !std::strcmp(argv[1], "--no-gpu");
    int nthreads = 8, ierr = 0;

    // host-only symbols
    {
        int rank = 4, dims[4] = {3, 4, 5, 6};
        check(tensor_size_by_shape_(rank, dims, ierr) == 360 && ierr == 0, "tensor_size_by_shape_");
        // ring: R(a,i,b,j) = W(k,b,c,j) T(a,i,c,k), labels a=1 i=2 b=3 j=4 k=5 c=6
        int dr = 4, lr = 4, rr = 4;
        int aces[12] = {1, 2, 3, 4, 5, 3, 6, 4, 1, 2, 6, 5};
        int ptrn[8];
        get_contraction_ptrn_(dr, lr, rr, aces, ptrn, ierr);
        const int want[8] = {-4, 3, -3, 4, 1, 2, -3, -1};
        check(ierr == 0 && !std::memcmp(ptrn, want, sizeof(want)), "get_contraction_ptrn_ (ring)");
    }

    // contraction R(a,i,b,j) = sum_{k,c} W(k,b,c,j) T(a,i,c,k), small uneven extents
    const int A = 5, I = 3, B = 4, J = 2, K = 3, C = 6;
    int wd[4] = {K, B, C, J}, td[4] = {A, I, C, K}, rd[4] = {A, I, B, J};
    std::vector<double> W = data(K * B * C * J, 1), T = data(A * I * C * K, 2);
    std::vector<double> R(A * I * B * J, 7.0);
    {
        int ptrn[8] = {-4, 3, -3, 4, 1, 2, -3, -1};
        int lr = 4, rr = 4, dr = 4;
        tensor_block_contract__(nthreads, ptrn, W.data(), lr, wd, T.data(), rr, td, R.data(), dr, rd,
                                ierr);
        if (no_gpu) {
            check(ierr != 0, "tensor_block_contract__ reports the missing device");
        } else {
            double err = 0.0, ref = 0.0;
            for (int a = 0; a < A; ++a)
                for (int i = 0; i < I; ++i)
                    for (int b = 0; b < B; ++b)
                        for (int j = 0; j < J; ++j) {
                            long double s = 0.0L;
                            for (int k = 0; k < K; ++k)
                                for (int c = 0; c < C; ++c)
                                    s += (long double)W[k + K * (b + B * (c + C * j))] *
                                         T[a + A * (i + I * (c + C * k))];
                            const double got = R[a + A * (i + I * (b + B * j))];
                            err = std::fmax(err, std::fabs(got - (double)s));
                            ref = std::fmax(ref, std::fabs((double)s));
                        }
            check(ierr == 0 && err <= 1e-13 * ref, "tensor_block_contract__ (assign semantics)");
        }
    }

    // permute: out(j,k,i) = in(i,j,k) for dims (4,5,6): dim_transp = [1, 3, 1, 2]
    {
        int rank = 3, dims[3] = {4, 5, 6}, tr[4] = {1, 3, 1, 2};
        std::vector<double> in = data(120, 3), out(120, 0.0);
        tensor_block_copy__(nthreads, rank, dims, tr, in.data(), out.data(), ierr);
        if (no_gpu) {
            check(ierr != 0, "tensor_block_copy__ reports the missing device");
        } else {
            bool ok = ierr == 0;
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 5; ++j)
                    for (int k = 0; k < 6; ++k)
                        ok = ok && out[j + 5 * (k + 6 * i)] == in[i + 4 * (j + 5 * k)];
            check(ok, "tensor_block_copy__");
        }
    }

    // slice and insert: a 3 x 2 x 4 slice at (1, 3, 2) of a 6 x 5 x 7 block
    {
        int rank = 3, te[3] = {6, 5, 7}, se[3] = {3, 2, 4}, eb[3] = {1, 3, 2};
        std::vector<double> tens = data(210, 4), slc(24, 0.0), ins = data(24, 5);
        tensor_block_slice__(nthreads, rank, tens.data(), te, slc.data(), se, eb, ierr);
        if (no_gpu) {
            check(ierr != 0, "tensor_block_slice__ reports the missing device");
        } else {
            bool ok = ierr == 0;
            for (int x = 0; x < 3; ++x)
                for (int y = 0; y < 2; ++y)
                    for (int z = 0; z < 4; ++z)
                        ok = ok && slc[x + 3 * (y + 2 * z)] ==
                                       tens[(x + 1) + 6 * ((y + 3) + 5 * (z + 2))];
            check(ok, "tensor_block_slice__");
            std::vector<double> t2 = tens;
            tensor_block_insert__(nthreads, rank, t2.data(), te, ins.data(), se, eb, ierr);
            ok = ierr == 0;
            for (int z = 0; z < 7; ++z)
                for (int y = 0; y < 5; ++y)
                    for (int x = 0; x < 6; ++x) {
                        const bool in = x >= 1 && x < 4 && y >= 3 && y < 5 && z >= 2 && z < 6;
                        const double want =
                            in ? ins[(x - 1) + 3 * ((y - 3) + 2 * (z - 2))] : tens[x + 6 * (y + 5 * z)];
                        ok = ok && t2[x + 6 * (y + 5 * z)] == want;
                    }
            check(ok, "tensor_block_insert__");
        }
        int bad = -1;
        tensor_block_slice__(nthreads, bad, tens.data(), te, slc.data(), se, eb, ierr);
        check(ierr == 1, "tensor_block_slice__ negative rank: ierr 1");
    }

    // init, scale, add, norm2 on a 7 x 9 block
    {
        int rank = 2, dims[2] = {7, 9};
        std::vector<double> x = data(63, 6), y = data(63, 7);
        double v = 2.5, f = 0.75;
        std::vector<double> t = x;
        tensor_block_init__(nthreads, t.data(), rank, dims, v, ierr);
        if (no_gpu) {
            check(ierr != 0, "tensor_block_init__ reports the missing device");
        } else {
            bool ok = ierr == 0;
            for (double e : t) ok = ok && e == 2.5;
            check(ok, "tensor_block_init__");
            t = x;
            tensor_block_scale__(nthreads, t.data(), rank, dims, f, ierr);
            ok = ierr == 0;
            for (int i = 0; i < 63; ++i) ok = ok && t[i] == x[i] * 0.75;
            check(ok, "tensor_block_scale__");
            t = x;
            tensor_block_add__(nthreads, rank, dims, t.data(), y.data(), f, ierr);
            ok = ierr == 0;
            for (int i = 0; i < 63; ++i) {
                volatile double p = y[i] * 0.75; // rounded product, then the sum
                ok = ok && t[i] == x[i] + p;
            }
            check(ok, "tensor_block_add__");
            long double s = 0.0L;
            for (double e : x) s += (long double)e * e;
            const double n2 = tensor_block_norm2__(nthreads, x.data(), rank, dims, ierr);
            check(ierr == 0 && std::fabs(n2 - (double)s) <= 1e-14 * (double)s, "tensor_block_norm2__");
        }
        int bad = -1;
        tensor_block_init__(nthreads, x.data(), bad, dims, v, ierr);
        check(ierr == -1, "tensor_block_init__ negative rank: ierr -1");
    }

    std::printf("%s\n", failures ? "FAILED" : "all passed");
    return failures ? 1 : 0;
}
