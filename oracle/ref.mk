# Compiles the reference's own C sources (read in place from
# /root/reference/cuxmss, never copied) into oracle/_ref/:
#   libxmssref.so       plain core   (xmss_core.c)
#   libxmssref_fast.so  BDS core     (xmss_core_fast.c)
#   vectors             the reference's KAT emitter (test/vectors.c)
# randombytes.c is replaced by ref_rng.c (seeded SplitMix64 stream).
# Needs gcc and the image's OpenSSL libcrypto (hash.c calls SHA256()).
# Test infrastructure only; the outputs are git-ignored.
REF ?= /root/reference/cuxmss
OUT := _ref
CC ?= gcc
CFLAGS := -O3 -fPIC -w
COMMON := params.c hash.c fips202.c hash_address.c wots.c xmss.c \
          xmss_commons.c utils.c sha256.c sha512.c
SRC := $(addprefix $(REF)/,$(COMMON))

all: $(OUT)/libxmssref.so $(OUT)/libxmssref_fast.so $(OUT)/vectors

$(OUT)/libxmssref.so: $(SRC) $(REF)/xmss_core.c ref_rng.c
	@mkdir -p $(OUT)
	$(CC) $(CFLAGS) -shared -o $@ $(SRC) $(REF)/xmss_core.c ref_rng.c -lcrypto

$(OUT)/libxmssref_fast.so: $(SRC) $(REF)/xmss_core_fast.c ref_rng.c
	@mkdir -p $(OUT)
	$(CC) $(CFLAGS) -shared -o $@ $(SRC) $(REF)/xmss_core_fast.c ref_rng.c -lcrypto

$(OUT)/vectors: $(SRC) $(REF)/xmss_core.c $(REF)/test/vectors.c ref_rng.c
	@mkdir -p $(OUT)
	$(CC) -O3 -w -o $@ $(SRC) $(REF)/xmss_core.c $(REF)/test/vectors.c ref_rng.c -lcrypto

.PHONY: all
