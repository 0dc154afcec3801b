/* Writes small particle-field HDF5 files with the HDF5 library itself, as
 * fixtures for the reader in flowunsteady_amd/h5.py: the flat root-group
 * datasets of a particle-field file (np, nt, t, X, Gamma, sigma, circulation,
 * vol, static, i, connectivity) for 5 particles with known values, in three
 * storage forms:
 *   pfield_contiguous.h5  default file format (earliest library version):
 *                         superblock 0, symbol-table root group, contiguous
 *   pfield_chunked.h5     the same with X and Gamma chunked (1 x 3) and the
 *                         1-D arrays chunked (2): version-1 B-tree chunk index
 *   pfield_latest.h5      library-version bounds (latest, latest):
 *                         superblock 3, version-2 object headers, link
 *                         messages, version-4 layouts
 * Values: X[k][c] = k + c/10, Gamma[k][c] = -(k + 1) * 10^-(c+1),
 * sigma[k] = 0.01 (k + 1), circulation[k] = 0.5 k, vol[k] = 1e-3 / (k + 1),
 * static[k] = k % 2, i[k] = k + 1, np = 5, nt = 12, t = 0.375,
 * connectivity = (1, 1, k) per particle.
 * Build and run: make_h5_fixtures.sh (h5cc from the HDF5 installation). */
#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>

#define NP 5

static void put(hid_t f, const char *name, hid_t type, int rank, const hsize_t *dims,
                const hsize_t *chunk, const void *data)
{
    hid_t sp = rank == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(rank, dims, NULL);
    hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
    if (chunk && rank > 0) H5Pset_chunk(dcpl, rank, chunk);
    hid_t d = H5Dcreate2(f, name, type, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT);
    if (d < 0 || H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) {
        fprintf(stderr, "writing %s failed\n", name);
        exit(1);
    }
    H5Dclose(d);
    H5Pclose(dcpl);
    H5Sclose(sp);
}

static void write_file(const char *path, int chunked, int latest)
{
    double X[NP][3], G[NP][3], sigma[NP], circ[NP], vol[NP], stat[NP], t = 0.375;
    long long idx[NP], conn[3 * NP], np_ = NP, nt = 12;
    for (int k = 0; k < NP; k++) {
        double p = 1.0;
        for (int c = 0; c < 3; c++) {
            p /= 10.0;
            X[k][c] = k + c / 10.0;
            G[k][c] = -(k + 1) * p;
        }
        sigma[k] = 0.01 * (k + 1);
        circ[k] = 0.5 * k;
        vol[k] = 1e-3 / (k + 1);
        stat[k] = k % 2;
        idx[k] = k + 1;
        conn[3 * k] = 1;
        conn[3 * k + 1] = 1;
        conn[3 * k + 2] = k;
    }
    hid_t fapl = H5Pcreate(H5P_FILE_ACCESS);
    if (latest) H5Pset_libver_bounds(fapl, H5F_LIBVER_LATEST, H5F_LIBVER_LATEST);
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, fapl);
    if (f < 0) {
        fprintf(stderr, "cannot create %s\n", path);
        exit(1);
    }
    const hsize_t d2[2] = {NP, 3}, c2[2] = {1, 3}, d1[1] = {NP}, c1[1] = {2};
    const hsize_t dc[1] = {3 * NP}, cc[1] = {4};
    put(f, "np", H5T_STD_I64LE, 0, NULL, NULL, &np_);
    put(f, "nt", H5T_STD_I64LE, 0, NULL, NULL, &nt);
    put(f, "t", H5T_IEEE_F64LE, 0, NULL, NULL, &t);
    put(f, "X", H5T_IEEE_F64LE, 2, d2, chunked ? c2 : NULL, X);
    put(f, "Gamma", H5T_IEEE_F64LE, 2, d2, chunked ? c2 : NULL, G);
    put(f, "sigma", H5T_IEEE_F64LE, 1, d1, chunked ? c1 : NULL, sigma);
    put(f, "circulation", H5T_IEEE_F64LE, 1, d1, chunked ? c1 : NULL, circ);
    put(f, "vol", H5T_IEEE_F64LE, 1, d1, chunked ? c1 : NULL, vol);
    put(f, "static", H5T_IEEE_F64LE, 1, d1, chunked ? c1 : NULL, stat);
    put(f, "i", H5T_STD_I64LE, 1, d1, chunked ? c1 : NULL, idx);
    put(f, "connectivity", H5T_STD_I64LE, 1, dc, chunked ? cc : NULL, conn);
    H5Fclose(f);
    H5Pclose(fapl);
}

int main(int argc, char **argv)
{
    const char *dir = argc > 1 ? argv[1] : ".";
    char p[4096];
    snprintf(p, sizeof p, "%s/pfield_contiguous.h5", dir);
    write_file(p, 0, 0);
    snprintf(p, sizeof p, "%s/pfield_chunked.h5", dir);
    write_file(p, 1, 0);
    snprintf(p, sizeof p, "%s/pfield_latest.h5", dir);
    write_file(p, 0, 1);
    return 0;
}
