// Writes golden random streams of libstdc++'s std::default_random_engine and
// std::bernoulli_distribution (the generator and distribution the DQMC
// reference draws its fields and Metropolis decisions from).  Built and run by
// gen_rng.py; the output rng_golden.json is committed.
#include <cstdio>
#include <random>
#include <string>
#include <vector>

static void hexbits(const std::vector<int>& b) {
    for (size_t i = 0; i < b.size(); i += 4) {
        int v = 0;
        for (int k = 0; k < 4; ++k) v |= (i + k < b.size() ? b[i + k] : 0) << k;
        printf("%x", v);
    }
}

int main() {
    const long long seeds[] = {0, 1, 42, 255, 12345, -7};
    printf("{\n");
    bool first = true;
    for (long long s : seeds) {
        if (!first) printf(",\n");
        first = false;
        printf("\"%lld\": {", s);
        // 1) 32768 draws of bernoulli(0.5)
        {
            std::default_random_engine e((int)s);
            std::bernoulli_distribution d(0.5);
            std::vector<int> b(32768);
            for (auto& x : b) x = d(e) ? 1 : 0;
            printf("\"p05_bits\": \"");
            hexbits(b);
            printf("\", ");
        }
        // 2) bernoulli with a sweep of probabilities p_k = (k % 101) / 100 (p=1 included)
        {
            std::default_random_engine e((int)s);
            std::vector<int> b(20000);
            for (size_t k = 0; k < b.size(); ++k) {
                std::bernoulli_distribution d((double)(k % 101) / 100.0);
                b[k] = d(e) ? 1 : 0;
            }
            printf("\"pvar_bits\": \"");
            hexbits(b);
            printf("\", ");
        }
        // 3) raw engine outputs and canonical doubles
        {
            std::default_random_engine e((int)s);
            printf("\"raw\": [");
            for (int k = 0; k < 16; ++k) printf("%s%lu", k ? ", " : "", (unsigned long)e());
            printf("], ");
            std::default_random_engine e2((int)s);
            printf("\"canonical\": [");
            for (int k = 0; k < 64; ++k) printf("%s\"%a\"", k ? ", " : "", std::generate_canonical<double, 53>(e2));
            printf("]");
        }
        printf("}");
    }
    printf("\n}\n");
    return 0;
}
