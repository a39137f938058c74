// One row of the per-window table (the reference's Window, src/Window.hpp:7-37).
#pragma once

#include <cstdlib>
#include <string>

class Window
{
public:
    std::string id;
    int start = 0;
    int end = 0;
    int totalKmers = 0;
    int observedKmers = 0;
    int variations = 0;
    int kmerDistance = 0;

    // Kept for interface compatibility; the reference never calls it.
    void addVariation(int gapSize, int kmerSize)
    {
        variations += 1;
        kmerDistance = gapSize < kmerSize ? gapSize : gapSize - (kmerSize - 1);
    }

    // One maximal run of `gapSize` missing k-mers: one variation, distance
    // gap-(k-1), or |gap-(k-1)+1| when that is not positive.
    void addVariationOriginal(int gapSize, int kmerSize)
    {
        variations += 1;
        int d = gapSize - (kmerSize - 1);
        if (d <= 0)
            d = std::abs(d + 1);
        kmerDistance += d;
    }
};
