// Writes the permutation that thrust::shuffle with a default-constructed
// thrust::default_random_engine applies to 0..m-1 (the call the reference
// makes on its face array, simp_cuda/src/cusimp_free.cu:937-941), through
// rocThrust on the device, for each m given on the command line.
// Output: for each m, the int32 m followed by m int32 values, to stdout.
#include <thrust/device_vector.h>
#include <thrust/host_vector.h>
#include <thrust/random.h>
#include <thrust/sequence.h>
#include <thrust/shuffle.h>

#include <cstdio>
#include <cstdlib>

int main(int argc, char** argv) {
    for (int a = 1; a < argc; ++a) {
        const int m = std::atoi(argv[a]);
        thrust::device_vector<int> v(m);
        thrust::sequence(v.begin(), v.end());
        thrust::default_random_engine rng;
        thrust::shuffle(v.begin(), v.end(), rng);
        thrust::host_vector<int> h = v;
        std::fwrite(&m, sizeof(int), 1, stdout);
        std::fwrite(h.data(), sizeof(int), (size_t)m, stdout);
    }
    return 0;
}
