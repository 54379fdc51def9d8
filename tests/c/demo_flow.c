/* Client -> server -> client round trip through key and ciphertext files, in the shape of the
 * reference's demo (test_bench/client/alice.c, test_bench/server/cloud.c,
 * test_bench/client/verify.c), built against libtfhe_mi355x.
 *   demo_flow client DIR X Y   key set (seed 314,1592,657), keys and 2x16 encrypted bits to DIR
 *   demo_flow server DIR       reads the cloud key and the bits, writes OR of each bit pair
 *   demo_flow verify DIR       reads the secret key and the answer, prints the 16-bit value */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tfhe_mi355x.h"

static FILE* open_in(const char* dir, const char* name, const char* mode) {
    char path[4096];
    snprintf(path, sizeof path, "%s/%s", dir, name);
    FILE* f = fopen(path, mode);
    if (!f) {
        fprintf(stderr, "cannot open %s\n", path);
        exit(2);
    }
    return f;
}

static int client(const char* dir, int x, int y) {
    TFheGateBootstrappingParameterSet* params = new_default_gate_bootstrapping_parameters(110);
    uint32_t seed[] = {314, 1592, 657};
    tfhe_random_generator_setSeed(seed, 3);
    TFheGateBootstrappingSecretKeySet* key = new_random_gate_bootstrapping_secret_keyset(params);
    FILE* f = open_in(dir, "secret.key", "wb");
    export_tfheGateBootstrappingSecretKeySet_toFile(f, key);
    fclose(f);
    f = open_in(dir, "cloud.key", "wb");
    export_tfheGateBootstrappingCloudKeySet_toFile(f, &key->cloud);
    fclose(f);
    LweSample* cx = new_gate_bootstrapping_ciphertext_array(16, params);
    LweSample* cy = new_gate_bootstrapping_ciphertext_array(16, params);
    for (int i = 0; i < 16; i++) bootsSymEncrypt(&cx[i], (x >> i) & 1, key);
    for (int i = 0; i < 16; i++) bootsSymEncrypt(&cy[i], (y >> i) & 1, key);
    f = open_in(dir, "cloud.data", "wb");
    for (int i = 0; i < 16; i++) export_gate_bootstrapping_ciphertext_toFile(f, &cx[i], params);
    for (int i = 0; i < 16; i++) export_gate_bootstrapping_ciphertext_toFile(f, &cy[i], params);
    fclose(f);
    delete_gate_bootstrapping_ciphertext_array(16, cy);
    delete_gate_bootstrapping_ciphertext_array(16, cx);
    delete_gate_bootstrapping_secret_keyset(key);
    delete_gate_bootstrapping_parameters(params);
    return 0;
}

static int server(const char* dir) {
    FILE* f = open_in(dir, "cloud.key", "rb");
    TFheGateBootstrappingCloudKeySet* bk = new_tfheGateBootstrappingCloudKeySet_fromFile(f);
    fclose(f);
    if (!bk) {
        fprintf(stderr, "cloud key: %s\n", tfhe_gpu_last_error());
        return 1;
    }
    const TFheGateBootstrappingParameterSet* params = bk->params;
    LweSample* cx = new_gate_bootstrapping_ciphertext_array(16, params);
    LweSample* cy = new_gate_bootstrapping_ciphertext_array(16, params);
    LweSample* res = new_gate_bootstrapping_ciphertext_array(16, params);
    f = open_in(dir, "cloud.data", "rb");
    for (int i = 0; i < 16; i++) import_gate_bootstrapping_ciphertext_fromFile(f, &cx[i], params);
    for (int i = 0; i < 16; i++) import_gate_bootstrapping_ciphertext_fromFile(f, &cy[i], params);
    fclose(f);
    for (int i = 0; i < 16; i++) bootsOR(&res[i], &cx[i], &cy[i], bk);
    f = open_in(dir, "answer.data", "wb");
    for (int i = 0; i < 16; i++) export_gate_bootstrapping_ciphertext_toFile(f, &res[i], params);
    fclose(f);
    delete_gate_bootstrapping_ciphertext_array(16, res);
    delete_gate_bootstrapping_ciphertext_array(16, cy);
    delete_gate_bootstrapping_ciphertext_array(16, cx);
    delete_gate_bootstrapping_cloud_keyset(bk);
    return 0;
}

static int verify(const char* dir) {
    FILE* f = open_in(dir, "secret.key", "rb");
    TFheGateBootstrappingSecretKeySet* key = new_tfheGateBootstrappingSecretKeySet_fromFile(f);
    fclose(f);
    if (!key) {
        fprintf(stderr, "secret key: %s\n", tfhe_gpu_last_error());
        return 1;
    }
    LweSample* ans = new_gate_bootstrapping_ciphertext_array(16, key->params);
    f = open_in(dir, "answer.data", "rb");
    for (int i = 0; i < 16; i++) import_gate_bootstrapping_ciphertext_fromFile(f, &ans[i], key->params);
    fclose(f);
    int v = 0;
    for (int i = 0; i < 16; i++) v |= bootsSymDecrypt(&ans[i], key) << i;
    printf("%d\n", v);
    delete_gate_bootstrapping_ciphertext_array(16, ans);
    delete_gate_bootstrapping_secret_keyset(key);
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 5 && !strcmp(argv[1], "client")) return client(argv[2], atoi(argv[3]), atoi(argv[4]));
    if (argc >= 3 && !strcmp(argv[1], "server")) return server(argv[2]);
    if (argc >= 3 && !strcmp(argv[1], "verify")) return verify(argv[2]);
    fprintf(stderr, "usage: demo_flow client DIR X Y | server DIR | verify DIR\n");
    return 2;
}
