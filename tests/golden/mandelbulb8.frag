#version 450 core
// Power-8 Mandelbulb distance estimator, 8 iterations, bailout radius 2.
float sdf(vec3 p) {
    p.xyz = p.xzy;
    vec3 z = p;
    float power = 8.0;
    float r, theta, phi;
    float dr = 1.0;
    for (int i = 0; i < 8; ++i) {
        r = length(z);
        if (r > 2.0) continue;
        theta = atan(z.y / z.x);
        phi = asin(z.z / r);
        dr = pow(r, power - 1.0) * dr * power + 1.0;
        r = pow(r, power);
        theta = theta * power;
        phi = phi * power;
        z = r * vec3(cos(theta) * cos(phi), sin(theta) * cos(phi), sin(phi)) + p;
    }
    return 0.5 * log(r) * r / dr - 0.003;
}

void main() {}
