// Float2D.hh -- 2D float array view with the reference's storage order
// (src/tools/Float2D.hh:56-95): a[x][y] = raw[x * rows + y], so one column
// (fixed x) is contiguous in y.
#ifndef SWE_HIP_FLOAT2D_HH
#define SWE_HIP_FLOAT2D_HH

class Float2D {
public:
    int getRows() const { return rows; }
    int getCols() const { return cols; }
    float* getRawPointer() const { return rawData; }
    void setRawPointer(float* p) { rawData = p; }
    inline float* operator[](int x) { return rawData + (long)rows * x; }
    inline const float* operator[](int x) const { return rawData + (long)rows * x; }

protected:
    Float2D() : cols(0), rows(0), rawData(nullptr) {}
    Float2D(int c, int r) : cols(c), rows(r), rawData(nullptr) {}
    ~Float2D() {}
    int cols;
    int rows;
    float* rawData;
};

#endif
